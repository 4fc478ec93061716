/* Test infrastructure: the dem_* entry points of include/mercurydpm_dem.h
 * served by the CPU oracle, so that drivers written against the host API
 * (mercurydpm_fork_amd/host) can be linked against the oracle and checked on a
 * machine without a GPU.  It is never part of the product: the product links
 * libmercurydpm_dem.so (HIP).  Entry points the oracle has no counterpart for
 * (statistics, profiling) report zeros. */
#include <string.h>

#include "dem_oracle.h"

#define O(ctx) ((oracle_ctx*)(ctx))

int dem_abi_version(void) { return DEM_ABI_VERSION; }
int dem_create(dem_ctx** out, const dem_config* cfg) { (void)cfg; return oracle_dem_create((oracle_ctx**)out); }
int dem_destroy(dem_ctx* ctx) { return oracle_dem_destroy(O(ctx)); }
const char* dem_last_error(const dem_ctx* ctx) { return ctx ? oracle_dem_last_error((const oracle_ctx*)ctx) : ""; }
int dem_set_species(dem_ctx* ctx, int32_t n, const dem_pair_params* t) { return oracle_dem_set_species(O(ctx), n, t); }
int dem_set_walls(dem_ctx* ctx, int32_t n, const dem_wall* w) { return oracle_dem_set_walls(O(ctx), n, w); }
int dem_set_wall_groups(dem_ctx* ctx, int32_t n, const int32_t* first, const int32_t* count)
{
    return oracle_dem_set_wall_groups(O(ctx), n, first, count);
}
int dem_set_periodic(dem_ctx* ctx, int32_t n, const dem_periodic* b) { return oracle_dem_set_periodic(O(ctx), n, b); }
int dem_set_gravity(dem_ctx* ctx, const double g[3]) { return oracle_dem_set_gravity(O(ctx), g); }
int dem_set_background_drag(dem_ctx* ctx, double d) { return oracle_dem_set_background_drag(O(ctx), d); }
int dem_set_timestep(dem_ctx* ctx, double dt) { return oracle_dem_set_timestep(O(ctx), dt); }
int dem_set_rotation(dem_ctx* ctx, int32_t r) { return oracle_dem_set_rotation(O(ctx), r); }
int dem_set_time(dem_ctx* ctx, double t, int64_t nt) { return oracle_dem_set_time(O(ctx), t, nt); }
int dem_upload_particles(dem_ctx* ctx, int64_t n, const uint32_t* id, const int32_t* species, const double* radius,
                         const double* inv_mass, const double* inv_inertia, const double* x, const double* v,
                         const double* w, const double* q, const double* f, const double* t)
{
    return oracle_dem_upload_particles(O(ctx), n, id, species, radius, inv_mass, inv_inertia, x, v, w, q, f, t);
}
int dem_upload_contacts(dem_ctx* ctx, int64_t n, const uint32_t* p_id, const uint32_t* i_id, const double* spring,
                        const uint32_t* ts)
{
    return oracle_dem_upload_contacts(O(ctx), n, p_id, i_id, spring, ts);
}
int dem_compute_forces(dem_ctx* ctx) { return oracle_dem_compute_forces(O(ctx)); }
int dem_step(dem_ctx* ctx, int64_t n) { return oracle_dem_step(O(ctx), n); }
int dem_step_begin(dem_ctx* ctx) { return oracle_dem_step_begin(O(ctx)); }
int dem_step_end(dem_ctx* ctx) { return oracle_dem_step_end(O(ctx)); }
int dem_set_kinematics(dem_ctx* ctx, int64_t n, const double* x, const double* v, const double* w, const double* q)
{
    return oracle_dem_set_kinematics(O(ctx), n, x, v, w, q);
}
int dem_add_particles(dem_ctx* ctx, int64_t k, const uint32_t* id, const int32_t* species, const double* radius,
                      const double* inv_mass, const double* inv_inertia, const double* x, const double* v,
                      const double* w, const double* q)
{
    return oracle_dem_add_particles(O(ctx), k, id, species, radius, inv_mass, inv_inertia, x, v, w, q);
}
int dem_find_in_region(dem_ctx* ctx, int32_t kind, const double* params, int64_t cap, int64_t* n_out,
                       int64_t* index_out)
{
    return oracle_dem_find_in_region(O(ctx), kind, params, cap, n_out, index_out);
}
int dem_remove_particles(dem_ctx* ctx, int64_t k, const int64_t* index)
{
    return oracle_dem_remove_particles(O(ctx), k, index);
}
int dem_check_particles(dem_ctx* ctx, int64_t k, const double* x, const double* radius, const int32_t* species,
                        const double* max_distance, int32_t* free_out)
{
    return oracle_dem_check_particles(O(ctx), k, x, radius, species, max_distance, free_out);
}
int dem_download_particles(dem_ctx* ctx, double* x, double* v, double* w, double* q, double* f, double* t)
{
    return oracle_dem_download_particles(O(ctx), x, v, w, q, f, t);
}
int dem_contact_count(dem_ctx* ctx, int64_t* n) { return oracle_dem_contact_count(O(ctx), n); }
int dem_download_contacts(dem_ctx* ctx, int64_t cap, int64_t* n_out, uint32_t* p_id, uint32_t* i_id, double* spring,
                          uint32_t* ts)
{
    return oracle_dem_download_contacts(O(ctx), cap, n_out, p_id, i_id, spring, ts);
}
int dem_download_interactions(dem_ctx* ctx, int64_t cap, int64_t* n_out, dem_interaction* out)
{
    return oracle_dem_download_interactions(O(ctx), cap, n_out, out);
}
int dem_cg_z(dem_ctx* ctx, const dem_cg_config* cfg, double* out) { return oracle_dem_cg_z(O(ctx), cfg, out); }
int dem_get_stats(dem_ctx* ctx, dem_stats* s)
{
    (void)ctx;
    memset(s, 0, sizeof *s);
    return DEM_OK;
}
int dem_set_profiling(dem_ctx* ctx, int32_t on)
{
    (void)ctx;
    (void)on;
    return DEM_OK;
}
