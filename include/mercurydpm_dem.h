/*
 * mercurydpm_dem.h -- C ABI of the MI355X (gfx950) DEM time-step engine.
 *
 * The engine executes MercuryDPM's per-timestep hot path on one GPU:
 *   integrateBeforeForceComputation -> periodic wrap -> periodic images ->
 *   broad phase (multilevel cell grid + Verlet lists) -> pair forces
 *   (LinearViscoelastic / HertzianViscoelastic normal laws, sliding /
 *   rolling / torsion / Mindlin friction) -> wall forces (InfiniteWall,
 *   CylindricalWall) -> external forces -> integrateAfterForceComputation ->
 *   eraseOldInteractions.
 *
 * Every entry point below replaces a piece of the reference's CPU path; the
 * comment of each names the reference function(s) (paths relative to the
 * MercuryDPM source tree).  A C++ host (mercurydpm_fork_amd/host, class
 * GpuMercury3D) keeps the DPMBase/Mercury3D, SpeciesHandler, WallHandler and
 * BoundaryHandler API and calls these functions from its computeOneTimeStep()
 * override, so a driver's computeAllForces()/hGridFindContacts path becomes a
 * drop-in on one GPU.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative DEM_ERR_* code on
 *    failure; dem_last_error() returns the message.  Nothing throws or exits
 *    across the ABI (the reference's logger(ERROR) calls std::exit,
 *    Kernel/Logger.cc:194; the host shim maps codes back to logger calls).
 *  - Host buffers are borrowed for the duration of a call only.  Vectors are
 *    interleaved per particle (x0,y0,z0,x1,...), i.e. the memory layout of an
 *    array of Vec3D; quaternions are (q0,q1,q2,q3) per particle.
 *  - Particles are identified by their 32-bit id (BaseObject::getId).  Their
 *    position in the upload arrays is the ParticleHandler index.  Ids must be
 *    distinct.  They increase with the index until a particle is removed
 *    (ParticleHandler::removeObject moves the last particle into the gap);
 *    the periodic-history rule of InteractionHandler::
 *    removeObjectKeepingPeriodics (InteractionHandler.cc:316-371) compares
 *    indices, the id order of DPMBase::computeInternalForce (DPMBase.cc:3130)
 *    compares ids, and the engine orders both by id: after removals, a pair
 *    whose periodic-image interaction meets its direct one keeps the record
 *    the id order selects.
 *  - All arithmetic is IEEE binary64 evaluated in the reference's operation
 *    order without contraction (no FMA), so contact sets, springs and
 *    per-contact forces are bit-identical to the CPU path on identical input.
 */
#ifndef MERCURYDPM_DEM_H
#define MERCURYDPM_DEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEM_ABI_VERSION 9

/* error codes */
#define DEM_OK 0
#define DEM_ERR_ARG (-1)        /* invalid argument / precondition violated */
#define DEM_ERR_STATE (-2)      /* call out of order (e.g. step before upload) */
#define DEM_ERR_DEVICE (-3)     /* HIP runtime error */
#define DEM_ERR_CAPACITY (-4)   /* neighbour list / hash table overflow */
#define DEM_ERR_UNSUPPORTED (-5)/* configuration outside the engine's scope */

/* Interaction partner key for wall contacts: DEM_WALL_KEY | wall index. */
#define DEM_WALL_KEY 0x80000000u

typedef struct dem_ctx dem_ctx;

/* Normal force laws (Kernel/Interactions/NormalForceInteractions). */
enum dem_normal_law {
    DEM_NORMAL_LINEAR_VISCOELASTIC = 0,  /* LinearViscoelasticInteraction.cc:105-132 */
    DEM_NORMAL_HERTZIAN_VISCOELASTIC = 1 /* HertzianViscoelasticInteraction.cc:106-141 */
};

/* Friction force laws (Kernel/Interactions/FrictionForceInteractions). */
enum dem_friction_law {
    DEM_FRICTION_NONE = 0,     /* EmptyFrictionInteraction */
    DEM_FRICTION_SLIDING = 1,  /* SlidingFrictionInteraction.cc:106-196 */
    DEM_FRICTION_FULL = 2,     /* FrictionInteraction.cc:108-222 (sliding+rolling+torsion) */
    DEM_FRICTION_MINDLIN = 3   /* MindlinInteraction.cc:279-344 */
};

/*
 * Resolved parameters of one (possibly mixed) species pair.  The host derives
 * them with the reference's species setters (e.g.
 * LinearViscoelasticNormalSpecies::setCollisionTimeAndRestitutionCoefficient,
 * LinearViscoelasticNormalSpecies.cc:225-245) and mixing rule
 * (BaseSpecies::average, BaseSpecies.cc:110-115); the engine only reads them.
 * Replaces the per-contact SpeciesHandler::getMixedObject lookup
 * (SpeciesHandler.cc:922-945).
 */
typedef struct dem_pair_params {
    int32_t normal_law;            /* enum dem_normal_law */
    int32_t friction_law;          /* enum dem_friction_law */
    int32_t constant_restitution;  /* BaseNormalForce::getConstantRestitution */
    int32_t reserved;
    double stiffness;              /* linear k */
    double dissipation;            /* linear gamma; Hertz: dimensionless eta */
    double elastic_modulus;        /* Hertz effective E* */
    double sliding_stiffness;      /* k_t (sliding / full) */
    double sliding_dissipation;    /* gamma_t; Mindlin: dimensionless */
    double sliding_mu;             /* getSlidingFrictionCoefficient */
    double sliding_mu_static;      /* getSlidingFrictionCoefficientStatic */
    double shear_modulus;          /* Mindlin effective G* */
    double rolling_stiffness, rolling_dissipation, rolling_mu, rolling_mu_static;
    double torsion_stiffness, torsion_dissipation, torsion_mu, torsion_mu_static;
    double interaction_distance;   /* adhesive range; must be 0 for these laws */
} dem_pair_params;

enum dem_wall_type {
    DEM_WALL_INFINITE = 0,     /* InfiniteWall.cc:151-175 */
    DEM_WALL_CYLINDER = 1,     /* CylindricalWall.cc:72-83 (axis = z through origin) */
    DEM_WALL_AXISYMMETRIC = 2  /* AxisymmetricIntersectionOfWalls.cc:93-128 */
};

/* Sub-walls of an AxisymmetricIntersectionOfWalls (at most this many). */
#define DEM_AXI_MAX_SUB 4

/*
 * One wall (WallHandler object).  For an InfiniteWall the host passes the
 * wall position p0 (BaseInteractable::getPosition), the three coefficients of
 * Quaternion::getDistance (Quaternion.cc:658-669: q00+q11-q22-q33, q12+q03,
 * q13-q02, with q12=2q1q2 etc.) and the normal Quaternion::getAxis
 * (Quaternion.cc:501-512) of the wall orientation, so the engine evaluates the
 * distance and normal with the reference's own rounding.
 *
 * An AxisymmetricIntersectionOfWalls (AxisymmetricIntersectionOfWalls.cc:
 * 93-128) is the body of revolution of an IntersectionOfWalls drawn in the
 * (radial, 0, axial) plane: position is its origin (getPosition), axis its
 * symmetry axis (getOrientation().getAxis()), and n_sub sub-walls (the
 * IntersectionOfWalls::addObject(normal, point) calls, in order,
 * IntersectionOfWalls.cc:138-148) are given like InfiniteWalls in that plane:
 * sub_position, sub_coef and sub_axis as position, dist_coef and axis above.
 * The engine derives the intersection lines and corners of the sub-walls
 * itself (IntersectionOfWalls::setPointsAndLines, IntersectionOfWalls.cc:
 * 351-443).
 */
typedef struct dem_wall {
    int32_t type;          /* enum dem_wall_type */
    int32_t species;       /* index into the species table */
    uint32_t id;           /* BaseObject id, reported in contact keys */
    int32_t n_sub;         /* AXISYMMETRIC: sub-walls (1..DEM_AXI_MAX_SUB); else 0 */
    double position[3];    /* InfiniteWall p0; AXISYMMETRIC: origin */
    double dist_coef[3];   /* InfiniteWall distance coefficients */
    double axis[3];        /* InfiniteWall normal (getAxis); AXISYMMETRIC: symmetry axis */
    double radius;         /* CylindricalWall radius */
    double sub_position[DEM_AXI_MAX_SUB][3];
    double sub_coef[DEM_AXI_MAX_SUB][3];
    double sub_axis[DEM_AXI_MAX_SUB][3];
} dem_wall;

/*
 * One PeriodicBoundary (PeriodicBoundary.cc:84-94), axis-aligned: normal_
 * = unit vector along `axis`, distanceLeft_/distanceRight_ as stored by
 * PeriodicBoundary::set, shift = (distanceRight_-distanceLeft_) (the non-zero
 * component of shift_).  Boundaries are applied in the given order, which is
 * the BoundaryHandler order (it decides how corner images are chained,
 * PeriodicBoundary.cc:349-364).
 */
typedef struct dem_periodic {
    int32_t axis;      /* 0 = x, 1 = y, 2 = z */
    int32_t reserved;
    double left;       /* distanceLeft_ */
    double right;      /* distanceRight_ */
    double shift;      /* distanceRight_ - distanceLeft_ */
} dem_periodic;

typedef struct dem_config {
    int32_t device;          /* HIP device ordinal */
    int32_t max_neighbors;   /* Verlet-list capacity per particle; 0 = 64 */
    double skin;             /* Verlet skin distance; <= 0 = 0.2 * min diameter */
    int32_t grid_levels;     /* cell-grid levels; 0 = automatic (<= 3) */
    int32_t precision;       /* DEM_PRECISION_FP64 (0) or DEM_PRECISION_MIXED */
    int32_t list_sweep;      /* Verlet-list build: DEM_SWEEP_PER_PARTICLE (0) or DEM_SWEEP_WAVEFRONT */
    int32_t reserved;        /* 0 */
} dem_config;

/* How the Verlet lists are built from the cell grid (the same lists either
 * way): every particle sweeps its own candidate cells, or the 64 particles of
 * a wavefront sweep the union of their cells together, each row of cells
 * loaded once into LDS (coalesced) and shared. */
enum dem_sweep { DEM_SWEEP_PER_PARTICLE = 0, DEM_SWEEP_WAVEFRONT = 1 };

/* Arithmetic of the contact forces.  FP64 follows the reference bit for bit
 * where the summation order allows; MIXED computes the contact geometry
 * (overlap, normal, lever arms, relative velocity) and the contact forces in
 * FP32 and keeps positions, velocities, the integration and the tangential
 * springs in FP64 (no reference counterpart: a precision option of this
 * engine). */
enum dem_precision { DEM_PRECISION_FP64 = 0, DEM_PRECISION_MIXED = 1 };

typedef struct dem_stats {
    int64_t n_particles;
    int64_t n_steps;              /* getNumberOfTimeSteps */
    int64_t n_rebuilds;           /* grid re-sorts + Verlet-list rebuilds */
    int64_t n_contacts;           /* stored interactions after the last step */
    int64_t n_wall_contacts;      /* of which particle-wall */
    int64_t n_list_entries;       /* Verlet-list entries (both directions) */
    int32_t max_list_length;      /* longest list at the last rebuild */
    int32_t grid_levels;
    double time;                  /* DPMBase::getTime */
    double skin;
    double cell_size[4];          /* per level */
    /* HIP-event times on the engine's stream, accumulated since the last
     * dem_set_profiling(ctx, 1) (zero when profiling is off) */
    int64_t n_timed_steps;
    double t_force_ms;            /* pair/wall/external force kernel */
    double t_kick_ms;             /* velocity-Verlet kick/drift kernels */
    double t_rebuild_ms;          /* grid sort + Verlet-list rebuild sequence */
    /* particles the device arrays hold: dem_add_particles appends in place
     * up to this count (beyond it, it re-uploads into larger arrays) */
    int64_t capacity;
} dem_stats;

/* Create/destroy a context bound to one GPU.  No counterpart object in the
 * reference: it owns the device state that Mercury3D's handlers hold on the
 * CPU (ParticleHandler, InteractionHandler, HGrid). */
int dem_create(dem_ctx** out, const dem_config* cfg);
int dem_destroy(dem_ctx* ctx);
const char* dem_last_error(const dem_ctx* ctx);
int dem_abi_version(void);

/* SpeciesHandler objects + mixed objects as an n x n table (row-major,
 * symmetric; entry [a*n+b] is getMixedObject(a,b)).  Replaces
 * SpeciesHandler::getMixedObject (SpeciesHandler.cc:922) and
 * InteractionHandler::getInteraction's species lookup (InteractionHandler.cc:149). */
int dem_set_species(dem_ctx* ctx, int32_t n_species, const dem_pair_params* table);

/* WallHandler contents (static walls).  Replaces the wall loop of
 * DPMBase::computeAllForces (DPMBase.cc:3434-3439) / computeWallForces
 * (DPMBase.cc:5451-5462).  Before dem_upload_particles it sets the walls;
 * afterwards it replaces them between steps (WallHandler::removeObject /
 * removeLastObject during a run, BaseHandler.h:472-505): at most as many
 * walls as at the upload; a wall that keeps its index and id keeps its
 * stored interactions, the interactions of every other index are deleted
 * (BaseInteractable::~BaseInteractable, BaseInteractable.cc:97-106). */
int dem_set_walls(dem_ctx* ctx, int32_t n_walls, const dem_wall* walls);

/* Per-species wall groups (no counterpart in the reference, where every wall
 * meets every particle; it lets many independent simulations -- each with its
 * own species and its own walls -- share one context, as a batched ensemble
 * of calibration heaps does): particles of species s meet only the walls
 * first[s] .. first[s] + count[s] - 1 (at most 32), in that order.  With
 * groups set there may be up to 65536 walls in all.  Before
 * dem_upload_particles it sets the groups (one per species of
 * dem_set_species); afterwards a group may only drop its last walls, whose
 * interactions are deleted, as WallHandler::removeLastObject deletes a
 * wall's interactions (BaseHandler.h:472-505, BaseInteractable.cc:97-106). */
int dem_set_wall_groups(dem_ctx* ctx, int32_t n_species, const int32_t* first, const int32_t* count);

/* BoundaryHandler periodic boundaries.  Replaces checkInteractionWithBoundaries
 * (DPMBase.cc:3272), checkAndDuplicatePeriodicParticles (DPMBase.cc:5043) and
 * removeDuplicatePeriodicParticles (DPMBase.cc:5000). */
int dem_set_periodic(dem_ctx* ctx, int32_t n_boundaries, const dem_periodic* b);

/* DPMBase::setGravity / setBackgroundDrag / setTimeStep / setRotation and the
 * time counters (time_, numberOfTimeSteps_).  Used by computeExternalForces
 * (DPMBase.cc:3168-3179) and the integrators. */
int dem_set_gravity(dem_ctx* ctx, const double g[3]);
int dem_set_background_drag(dem_ctx* ctx, double drag);
int dem_set_timestep(dem_ctx* ctx, double dt);
int dem_set_rotation(dem_ctx* ctx, int32_t rotation);
int dem_set_time(dem_ctx* ctx, double time, int64_t n_time_steps);

/*
 * Upload the ParticleHandler (real particles only; no periodic ghosts) in
 * handler order.  inv_inertia holds MatrixSymmetric3D (XX,XY,XZ,YY,YZ,ZZ) per
 * particle; spheres must be isotropic (BaseParticle::computeMass,
 * BaseParticle.cc:874-883).  f and t (force/torque of the previous force pass)
 * may be NULL (zero).  n may be 0: an empty system, to be filled by
 * dem_add_particles.  Replaces ParticleHandler::copyAndAddObject for the
 * device copy; invalidates all stored contacts (upload them afterwards).
 */
int dem_upload_particles(dem_ctx* ctx, int64_t n, const uint32_t* id, const int32_t* species,
                         const double* radius, const double* inv_mass, const double* inv_inertia,
                         const double* x, const double* v, const double* w, const double* q,
                         const double* f, const double* t);

/*
 * Upload the InteractionHandler contents (history of particle-particle and
 * particle-wall interactions).  p_id/i_id are BaseInteraction::getP()/getI()
 * ids (i_id = DEM_WALL_KEY|wall index for a wall); for particle pairs p_id
 * must be the lower id.  spring holds 9 doubles per contact: sliding spring,
 * rolling spring, torsion spring (zeros where unused); time_stamp is
 * BaseInteraction::getTimeStamp.
 */
int dem_upload_contacts(dem_ctx* ctx, int64_t n, const uint32_t* p_id, const uint32_t* i_id,
                        const double* spring, const uint32_t* time_stamp);

/* The force pass of DPMBase::initialiseSolve (DPMBase.cc:4260-4275):
 * images, broad phase, computeAllForces, remove images.  No integration, no
 * erase of old interactions. */
int dem_compute_forces(dem_ctx* ctx);

/* n_steps x DPMBase::computeOneTimeStep (DPMBase.cc:4315-4389) without
 * output: integrateBeforeForceComputation, checkInteractionWithBoundaries,
 * checkAndDuplicatePeriodicParticles, hGridActionsBeforeTimeStep,
 * computeAllForces, removeDuplicatePeriodicParticles,
 * integrateAfterForceComputation, eraseOldInteractions, time_ += dt. */
int dem_step(dem_ctx* ctx, int64_t n_steps);

/* One DPMBase::computeOneTimeStep split where the reference calls
 * actionsBeforeTimeStep (DPMBase.cc:4347), so that a host can prescribe
 * particle motion there as drivers do:
 *   dem_step_begin: integrateBeforeForceComputation and the periodic wrap of
 *                   checkInteractionWithBoundaries;
 *   dem_step_end:   checkAndDuplicatePeriodicParticles, hGridActions-
 *                   BeforeTimeStep, computeAllForces, removeDuplicate-
 *                   PeriodicParticles, integrateAfterForceComputation,
 *                   eraseOldInteractions, time_ += dt.
 * dem_step_begin + dem_step_end gives the same state as dem_step(ctx, 1). */
int dem_step_begin(dem_ctx* ctx);
int dem_step_end(dem_ctx* ctx);

/* BaseInteractable::setPosition / setVelocity / setAngularVelocity /
 * setOrientation for all n particles, arrays in handler order (x, v, w:
 * 3 doubles per particle, q: 4); a NULL array leaves that quantity as it is.
 * The stored interactions are kept, and the displacement counts towards the
 * Verlet-list rebuild like any other motion. */
int dem_set_kinematics(dem_ctx* ctx, int64_t n, const double* x, const double* v, const double* w,
                       const double* q);

/* ParticleHandler::copyAndAddObject during a run, as the insertion
 * boundaries do it (InsertionBoundary::checkBoundaryBeforeTimeStep,
 * InsertionBoundary.cc:184-310): append k particles after the resident ones
 * (handler indices n .. n+k-1).  Ids must increase and exceed every resident
 * id; arrays as dem_upload_particles (w and q may be NULL).  The resident
 * state and the stored interactions are kept.  Allowed between
 * dem_step_begin and dem_step_end, where the reference's boundaries insert
 * (DPMBase.cc:4348-4351): the new particles take part in that step's force
 * pass and closing half-kick.  An empty system (dem_upload_particles with
 * n = 0) is allowed: its steps only advance the clock.
 * Cost: when the particles fit the resident arrays (n + k <= capacity of
 * dem_get_stats) and each radius is within the largest radius of its grid
 * level, only the k new particles are copied; otherwise the whole state and
 * its interactions go through the host once, into arrays with a quarter more
 * room for the appends that follow. */
int dem_add_particles(dem_ctx* ctx, int64_t k, const uint32_t* id, const int32_t* species, const double* radius,
                      const double* inv_mass, const double* inv_inertia, const double* x, const double* v,
                      const double* w, const double* q);

/* ParticleHandler::removeObject (ParticleHandler.cc:410-426 with
 * BaseHandler::removeObject, BaseHandler.h:472-505) for index[0], index[1],
 * ... in turn: the particle at that handler index is deleted together with
 * its interactions, and the last particle takes its index.  Each index
 * refers to the handler as the previous removals left it.  Allowed between
 * steps and between dem_step_begin and dem_step_end, where deletion
 * boundaries act (DPMBase::checkInteractionWithBoundaries, DPMBase.cc:
 * 3272-3300).  Cost: the state stays on the device (only the k indices and
 * the k deleted ids cross PCIe), plus one grid and list rebuild over the
 * kept particles.  The ids of deleted particles are not given out again
 * while particles remain: later appends still need ids above every id the
 * context has held. */
int dem_remove_particles(dem_ctx* ctx, int64_t k, const int64_t* index);

/* The particles a deletion boundary removes, found on the device: the
 * handler indices (increasing; n_out = how many, of which at most capacity
 * are written) of the particles
 *   DEM_REGION_HALF_SPACE:  params = normal (3), distance: distance -
 *     dot(x, normal) < 0 (DeletionBoundary::getDistance, DeletionBoundary.cc:
 *     109-112);
 *   DEM_REGION_BOX_INTERIOR: params = min (3), max (3): strictly inside
 *     (CubeDeletionBoundary::getDistance, CubeDeletionBoundary.cc:60-69). */
enum dem_region { DEM_REGION_HALF_SPACE = 0, DEM_REGION_BOX_INTERIOR = 1 };
int dem_find_in_region(dem_ctx* ctx, int32_t kind, const double* params, int64_t capacity, int64_t* n_out,
                       int64_t* index_out);

/* DPMBase::checkParticleForInteraction (DPMBase.cc:4819-4905 with
 * MercuryBase::checkParticleForInteractionLocal, MercuryBase.cc:622-650) for
 * k candidate spheres against the resident particles (at their current
 * positions) and the walls: free_out[j] = 1 when candidate j (x: 3 doubles,
 * radius, species index) touches neither, 0 otherwise.  Near a periodic
 * boundary the candidate's images are tested as well; max_distance[j] is the
 * reference's image range, the candidate's getMaxInteractionRadius plus that
 * of the largest particle in the caller's handler, or a negative value when
 * that handler is empty (no images).  At most 65535 candidates per call. */
int dem_check_particles(dem_ctx* ctx, int64_t k, const double* x, const double* radius, const int32_t* species,
                        const double* max_distance, int32_t* free_out);

/* dem_check_particles without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The candidates are staged in a static array of the library, one per GPU,
 * sized for DEM_CHECK_STATIC_MAX candidates (DEM_ERR_CAPACITY above that).
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * race.  Same arguments, checks and results as dem_check_particles. */
#define DEM_CHECK_STATIC_MAX 4096
int dem_check_particles_static(dem_ctx* ctx, int64_t k, const double* x, const double* radius,
                               const int32_t* species, const double* max_distance, int32_t* free_out);

/* Download state in handler order (any pointer may be NULL). */
int dem_download_particles(dem_ctx* ctx, double* x, double* v, double* w, double* q,
                           double* f, double* t);

/* Number of stored interactions (InteractionHandler::getNumberOfObjects). */
int dem_contact_count(dem_ctx* ctx, int64_t* n);

/* Download the stored interactions, sorted by (p_id, i_id).  capacity is the
 * length of the arrays (spring: 9*capacity doubles).  n_out receives the
 * number written. */
int dem_download_contacts(dem_ctx* ctx, int64_t capacity, int64_t* n_out, uint32_t* p_id,
                          uint32_t* i_id, double* spring, uint32_t* time_stamp);

int dem_get_stats(dem_ctx* ctx, dem_stats* stats);

/* A stored interaction with the state BaseInteraction::writeToFStat
 * (BaseInteraction.cc:487-531) and getElasticEnergy read: the force on P and
 * the geometry of the last force pass.  `current` is 0 for an interaction
 * that pass did not evaluate (kept only by the time-stamp rules of
 * eraseOldInteractions); its force is then reported as zero. */
typedef struct dem_interaction {
    uint32_t p_id, i_id;  /* as dem_download_contacts (i_id: DEM_WALL_KEY | wall index for a wall) */
    uint32_t time_stamp;
    uint32_t current;
    double force[3];      /* BaseInteraction::getForce, on P */
    double tangential[3]; /* getTangentialForce: the friction part of force */
    double contact[3];    /* getContactPoint */
    double normal[3];     /* getNormal: from I towards P */
    double overlap, distance;
    double spring[9];     /* sliding, rolling, torsion springs */
} dem_interaction;

/* The stored interactions sorted by (p_id, i_id); capacity as in
 * dem_download_contacts. */
int dem_download_interactions(dem_ctx* ctx, int64_t capacity, int64_t* n_out, dem_interaction* out);

/* ---- coarse graining ----------------------------------------------------
 * CG<CGCoordinates::Z, CGFunctions::Gauss, CGFields::StandardFields>
 * evaluated on the current state, as cgHandler.evaluate() does in
 * DPMBase::writeOutputFiles (DPMBase.cc:4084): CG::evaluate (CG.hcc:248-254)
 * with evaluateParticle (CG.hcc:303-376) over the non-fixed particles and
 * evaluateContact (CG.hcc:445-527) over the stored interactions, the Gauss
 * function and its line integral (Gauss.hcc:79-241, Coordinates/Z.cc,
 * Base_X_Y_Z.cc:32-49), StandardFields::setFields (StandardFields.cc:232-264)
 * and the volume average over the x-y extent (CG.hcc:199-207).  Points are
 * spaced evenly in z (Coordinates/Z.h:126-138).  The contact forces are those
 * of the last force pass (the last step of a dem_step call, dem_step_end or
 * dem_compute_forces). */
typedef struct dem_cg_config {
    int32_t n;                   /* evaluation points in z (CG::setNZ) */
    int32_t particle_dimensions; /* 3 (default, 0) or 2: volume of a particle */
    double width;                /* Gauss width (CG::setWidth); cutoff 3 x width */
    double min[3], max[3];       /* CG domain; min[0..1] and max[0..1] give the averaged area */
} dem_cg_config;

/* Values per point, in this order (the columns of the reference's .stat
 * file after time and z): VolumeFraction, Density, Momentum (3),
 * MomentumFlux (xx xy xz yy yz zz), ContactStress (xx xy xz yx yy yz zx zy
 * zz), InteractionForce (3), ParticleSize (6). */
#define DEM_CG_FIELDS 29

/* out: n x DEM_CG_FIELDS doubles. */
int dem_cg_z(dem_ctx* ctx, const dem_cg_config* cfg, double* out);

/* Per particle, in handler order, 3 doubles: the kinetic energy 1/2 m v^2
 * (zero for fixed particles; DPMBase::getKineticEnergy, DPMBase.cc:
 * 1542-1553), the elastic energy of the stored interactions the particle
 * holds as P -- the pairs where it has the lower id, and its wall contacts --
 * (BaseInteraction::getElasticEnergy of each; DPMBase::getElasticEnergy,
 * DPMBase.cc:1528-1537, is their sum over all particles), and its height z,
 * at the state of the last force pass.  Lets a driver evaluate the
 * reference's kinetic/elastic-energy criteria (CalibrationHeap.cpp:108-140)
 * per sub-simulation without downloading the interactions. */
int dem_particle_energies(dem_ctx* ctx, double* out);

/* Per-phase HIP-event timing of dem_step (no counterpart in the reference,
 * whose only timer is the wall clock of Time::toc, MercuryTime.h:70-73).
 * Enabling resets the accumulated times. */
int dem_set_profiling(dem_ctx* ctx, int32_t on);

#ifdef __cplusplus
}
#endif

#endif /* MERCURYDPM_DEM_H */
