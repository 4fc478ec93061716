/*
 * dem_oracle.h -- CPU oracle of the DEM time-step engine (test infrastructure
 * only; see dem_oracle.c).  Same call sequence and argument meaning as the
 * engine's C ABI in include/mercurydpm_dem.h, with an oracle_ prefix.
 */
#ifndef MERCURYDPM_DEM_ORACLE_H
#define MERCURYDPM_DEM_ORACLE_H

#include "../include/mercurydpm_dem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

int oracle_dem_create(oracle_ctx** out);
int oracle_dem_destroy(oracle_ctx* c);
const char* oracle_dem_last_error(const oracle_ctx* c);
int oracle_dem_set_species(oracle_ctx* c, int32_t n, const dem_pair_params* table);
int oracle_dem_set_walls(oracle_ctx* c, int32_t n, const dem_wall* walls);
int oracle_dem_set_wall_groups(oracle_ctx* c, int32_t n, const int32_t* first, const int32_t* count);
int oracle_dem_set_periodic(oracle_ctx* c, int32_t n, const dem_periodic* b);
int oracle_dem_set_gravity(oracle_ctx* c, const double g[3]);
int oracle_dem_set_background_drag(oracle_ctx* c, double drag);
int oracle_dem_set_timestep(oracle_ctx* c, double dt);
int oracle_dem_set_rotation(oracle_ctx* c, int32_t rotation);
int oracle_dem_set_time(oracle_ctx* c, double time, int64_t nt);
int oracle_dem_upload_particles(oracle_ctx* c, int64_t n, const uint32_t* id, const int32_t* species,
                                const double* radius, const double* inv_mass, const double* inv_inertia,
                                const double* x, const double* v, const double* w, const double* q,
                                const double* f, const double* t);
int oracle_dem_upload_contacts(oracle_ctx* c, int64_t n, const uint32_t* p_id, const uint32_t* i_id,
                               const double* spring, const uint32_t* ts);
int oracle_dem_compute_forces(oracle_ctx* c);
int oracle_dem_step(oracle_ctx* c, int64_t n);
int oracle_dem_step_begin(oracle_ctx* c);
int oracle_dem_step_end(oracle_ctx* c);
int oracle_dem_set_kinematics(oracle_ctx* c, int64_t n, const double* x, const double* v, const double* w,
                              const double* q);
int oracle_dem_download_particles(oracle_ctx* c, double* x, double* v, double* w, double* q, double* f,
                                  double* t);
int oracle_dem_add_particles(oracle_ctx* c, int64_t k, const uint32_t* id, const int32_t* species,
                             const double* radius, const double* inv_mass, const double* inv_inertia,
                             const double* x, const double* v, const double* w, const double* q);
int oracle_dem_find_in_region(oracle_ctx* c, int32_t kind, const double* params, int64_t cap, int64_t* n_out,
                              int64_t* index_out);
int oracle_dem_remove_particles(oracle_ctx* c, int64_t k, const int64_t* index);
int oracle_dem_check_particles(oracle_ctx* c, int64_t k, const double* x, const double* radius,
                               const int32_t* species, const double* max_distance, int32_t* free_out);
int oracle_dem_contact_count(oracle_ctx* c, int64_t* n);
int oracle_dem_download_contacts(oracle_ctx* c, int64_t cap, int64_t* n_out, uint32_t* p_id, uint32_t* i_id,
                                 double* spring, uint32_t* ts);
int oracle_dem_download_interactions(oracle_ctx* c, int64_t cap, int64_t* n_out, dem_interaction* out);
int oracle_dem_get_time(oracle_ctx* c, double* time, int64_t* nt);
int oracle_dem_force_scale(oracle_ctx* c, double* fs, double* ts);
int oracle_dem_cg_z(oracle_ctx* c, const dem_cg_config* cfg, double* out);
int oracle_dem_set_threads(int32_t n);

#ifdef __cplusplus
}
#endif

#endif
