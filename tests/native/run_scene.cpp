// A C++ caller of the C ABI (include/ndipc.h, include/ndipc_io.h) that does what the
// reference's main -> Simulation::run does for this path (nexdynipc/src/main.cpp,
// nexdynipc/src/App/Simulation.cpp:14-112): load a JSON scene, configure the solver as
// Simulation::run does, step until max_time (or --steps), and write the state and KPI
// CSV of every frame. Built by __graft_entry__.build() into tests/native/_build/.
//
//   run_scene <scene.json> [--steps N] [--out DIR] [--name NAME] [--euler]
//
// The device path integrates with implicit Euler; a scene asking for another integrator
// runs only with --euler. Scenes with entries the device path does not run (other joint
// types, drives) are refused. Exit status: 0 done, 1 usage, 2 load or device error,
// 3 refused.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ndipc.h"
#include "ndipc_io.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: run_scene <scene.json> [--steps N] [--out DIR] [--name NAME] [--euler]\n");
        return 1;
    }
    const char* path = argv[1];
    int steps = -1;
    std::string out_dir, out_name;
    bool euler = false;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--steps" && i + 1 < argc) steps = std::atoi(argv[++i]);
        else if (a == "--out" && i + 1 < argc) out_dir = argv[++i];
        else if (a == "--name" && i + 1 < argc) out_name = argv[++i];
        else if (a == "--euler") euler = true;
        else {
            std::fprintf(stderr, "run_scene: unknown argument %s\n", a.c_str());
            return 1;
        }
    }
    ndipc_scene_file* sf = nullptr;
    if (ndipc_scene_load(path, &sf) != 0) {
        std::fprintf(stderr, "run_scene: %s\n", ndipc_io_last_error());
        return 2;
    }
    ndipc_sim_settings set;
    ndipc_scene_settings(sf, &set);
    const int n_unsup = ndipc_scene_unsupported(sf, -1, nullptr, 0);
    if (n_unsup > 0) {
        char buf[512];
        for (int k = 0; k < n_unsup; ++k) {
            ndipc_scene_unsupported(sf, k, buf, sizeof buf);
            std::fprintf(stderr, "run_scene: not run by the device path: %s\n", buf);
        }
        ndipc_scene_free(sf);
        return 3;
    }
    if (std::strcmp(set.integrator_type, "ImplicitEuler") != 0 && !euler) {
        std::fprintf(stderr, "run_scene: the scene asks for %s; the device path integrates with implicit Euler "
                             "(pass --euler)\n", set.integrator_type);
        ndipc_scene_free(sf);
        return 3;
    }
    ndipc_scene_desc desc;
    ndipc_params params;
    ndipc_scene_describe(sf, &desc);
    ndipc_scene_params(sf, &params);
    const int n = desc.n_bodies;
    std::vector<double> p(3 * n), q(4 * n), v(3 * n), w(3 * n), al(3 * n, 0.0), aa(3 * n, 0.0);
    ndipc_scene_state(sf, p.data(), q.data(), v.data(), w.data());

    ndipc_ctx* ctx = nullptr;
    int rc = ndipc_create(0, &ctx);
    if (rc != 0) {
        std::fprintf(stderr, "run_scene: ndipc_create failed (%d)\n", rc);
        ndipc_scene_free(sf);
        return 2;
    }
    auto die = [&](const char* what) {
        std::fprintf(stderr, "run_scene: %s: %s\n", what, ndipc_last_error(ctx));
        ndipc_destroy(ctx);
        ndipc_scene_free(sf);
        return 2;
    };
    if (ndipc_upload_scenes(ctx, &desc) != 0) return die("upload");
    if (ndipc_configure(ctx, &params) != 0) return die("configure");
    if (ndipc_set_state(ctx, p.data(), q.data(), v.data(), w.data(), al.data(), aa.data()) != 0)
        return die("set_state");

    ndipc_exporter* ex = nullptr;
    const std::string dir = out_dir.empty() ? set.output_dir : out_dir;
    const std::string name = out_name.empty() ? set.output_name : out_name;
    if (ndipc_export_open(dir.c_str(), name.c_str(), &ex) != 0) {
        std::fprintf(stderr, "run_scene: %s\n", ndipc_io_last_error());
        ndipc_destroy(ctx);
        ndipc_scene_free(sf);
        return 2;
    }
    double t = 0.0;
    int frame = 0;
    ndipc_export_frame(ex, sf, frame, t, p.data(), q.data(), v.data(), w.data(), al.data(), aa.data());
    ndipc_export_kpi(ex, frame, t, 0.0, 0.0, 0.0, 0.0);
    ++frame;
    const double dt = set.dt;
    int done = 0;
    double res[3];
    while (steps >= 0 ? done < steps : t < set.max_time) {  // Simulation.cpp:71-90
        if (ndipc_step(ctx, &dt, 1) != 0) return die("step");
        t += dt;
        ++done;
        if (ndipc_get_state(ctx, p.data(), q.data(), v.data(), w.data(), al.data(), aa.data()) != 0)
            return die("get_state");
        if (ndipc_last_residuals(ctx, res) != 0) return die("last_residuals");
        ndipc_export_frame(ex, sf, frame, t, p.data(), q.data(), v.data(), w.data(), al.data(), aa.data());
        // V_w and T_sat belong to drives, which this path does not run
        ndipc_export_kpi(ex, frame, t, 0.0, 0.0, res[1], res[0]);
        ++frame;
    }
    ndipc_export_close(ex);
    std::printf("run_scene: %d steps of %d bodies, t = %.17g\n", done, n, t);
    ndipc_destroy(ctx);
    ndipc_scene_free(sf);
    return 0;
}
