/* nexdynipc_amd — scene and state I/O around the C ABI of include/ndipc.h.
 *
 * Host code (no device work): the reference's JSON scene format and OBJ meshes
 * (nexdynipc/src/App/SceneLoader.cpp:40-109,460-707), the settings Simulation::run
 * applies to IPCSolver (nexdynipc/src/App/Simulation.cpp:14-56), and the state / KPI
 * CSV of StateExporter (nexdynipc/src/App/StateExporter.cpp:13-76). A loaded scene
 * describes itself as one ndipc_scene_desc + ndipc_params upload; a caller then runs it
 * with ndipc_create / ndipc_upload_scenes / ndipc_configure / ndipc_set_state /
 * ndipc_step and writes frames with the exporter.
 *
 * Return values: 0 on success, < 0 on error (ndipc_io_last_error() names it).
 */
#ifndef NDIPC_IO_H
#define NDIPC_IO_H
#include <stdint.h>

#include "ndipc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SimulationConfig (nexdynipc/include/App/Simulation.h:19-53) after the JSON's
 * "settings" and "interactions.contact" blocks (SceneLoader.cpp:599-707), with the
 * reference's defaults for absent keys. */
typedef struct ndipc_sim_settings {
    double dt;
    double max_time;
    double newmark_beta;
    double newmark_gamma;
    double joint_stiffness;          /* given to joints left at the default 1e8 */
    int32_t alm_max_iters;
    int32_t kpi_gate_enabled;
    double alm_constraint_tolerance;
    double alm_dual_tolerance;
    double alm_hardening_trigger;
    double alm_hardening_ratio;
    double kpi_v_w_max;
    double kpi_t_sat_max;
    double kpi_r_dual_max;
    int32_t contact_enabled;
    int32_t n_contact_pairs;         /* pairs read from interactions.contact.pairs */
    double contact_thickness;        /* barrier dhat */
    double contact_stiffness;        /* > 0: fixed stiffness, adaptive barrier off */
    double contact_friction;         /* mean friction of the enabled pairs that give one */
    char integrator_type[32];        /* "ImplicitEuler" or "ImplicitNewmark" */
    char output_dir[512];
    char output_name[256];
} ndipc_sim_settings;

typedef struct ndipc_scene_file ndipc_scene_file;

/* Last error of this thread's I/O calls. */
const char* ndipc_io_last_error(void);

/* SceneLoader::load(filename, world, config) for a ".json" scene: bodies (id, name, mass,
 * position, inertia, OBJ mesh relative to the scene's folder, else to the working
 * directory), "joints" then "constraints" (duplicates by signature skipped, "revolute"
 * read as "hinge"), "settings" and "interactions.contact". */
int32_t ndipc_scene_load(const char* json_path, ndipc_scene_file** out);
void ndipc_scene_free(ndipc_scene_file* f);
int32_t ndipc_scene_settings(const ndipc_scene_file* f, ndipc_sim_settings* out);
/* The scene as one upload (n_scenes = 1). The arrays belong to f (valid until freed).
 * Joint bodies are world indices by body id (IPCSolver.cpp:966-1066: an unknown id maps
 * to index 0, as the reference's map lookup does). */
int32_t ndipc_scene_describe(const ndipc_scene_file* f, ndipc_scene_desc* out);
/* The IPCSolver settings Simulation::run applies (Simulation.cpp:20-56). */
int32_t ndipc_scene_params(const ndipc_scene_file* f, ndipc_params* out);
/* Initial state: p[3n], q[4n] (w, x, y, z), v[3n], w[3n] (the loader reads positions
 * only: identity orientations, zero velocities). */
int32_t ndipc_scene_state(const ndipc_scene_file* f, double* p, double* q, double* v, double* w);
int32_t ndipc_scene_body_name(const ndipc_scene_file* f, int32_t body, char* out, int32_t cap);
/* Entries the device path does not run (joint types other than fixed and hinge, drives,
 * springs, limits, a body whose OBJ could not be read): the count, and with out the
 * description of entry k. */
int32_t ndipc_scene_unsupported(const ndipc_scene_file* f, int32_t k, char* out, int32_t cap);

/* OBJ mesh (SceneLoader.cpp:40-109): "v x y z" and "f a b c ..." lines (1-based indices,
 * "a/b/c" tokens, polygons fan-triangulated). Counts in *nv, *nf; vertices and faces are
 * written when the pointers are not NULL and the capacities suffice. */
int32_t ndipc_obj_load(const char* path, int32_t* nv, double* verts, int32_t v_cap, int32_t* nf, int32_t* faces,
                       int32_t f_cap);

/* StateExporter: <dir>/<name>.csv (frame_idx,time,body_id,body_name,pos,quat x y z w,v,a,
 * w,alpha) and <dir>/<name>_kpi.csv, the reference's columns; numbers written with %.17g
 * (round-trip exact; the reference's ostream writes 6 significant digits). Frame 0
 * creates the files with their headers, later frames append. */
typedef struct ndipc_exporter ndipc_exporter;
int32_t ndipc_export_open(const char* dir, const char* name, ndipc_exporter** out);
int32_t ndipc_export_frame(ndipc_exporter* x, const ndipc_scene_file* f, int32_t frame, double time, const double* p,
                           const double* q, const double* v, const double* w, const double* al, const double* aa);
int32_t ndipc_export_kpi(ndipc_exporter* x, int32_t frame, double time, double v_w, double t_sat, double r_dual,
                         double max_constraint_violation);
void ndipc_export_close(ndipc_exporter* x);

#ifdef __cplusplus
}
#endif
#endif
