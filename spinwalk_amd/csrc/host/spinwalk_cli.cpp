// spinwalk_cli.cpp -- `spinwalk sim -c config.ini ...` on the MI355X engine.
//
// The command-line surface of the reference's sim subcommand and global flags
// (src/spinwalk.cpp:45-56): -c/--configs (one or more files), -p/--use_cpu,
// -d/--device, -g/--gpu_info, -v/--version.  -p is refused: this engine has no CPU
// path.  Extra sim options select the engine's modes: --rng philox|replay,
// --precision f32|kahan, --field f32|f16|bf16, --no-spatial-order.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/spinwalk_host.h"

extern "C" int sw_run_config_ex(sw_engine *eng, const char *path, int32_t rng, int32_t precision,
                                int32_t field_format, int32_t spatial_order);

static void usage() {
    std::printf(
        "Usage: spinwalk [-v] [-g] sim -c CONFIG... [-d DEVICE] [--rng philox|replay]\n"
        "                [--precision f32|kahan] [--field f32|f16|bf16] [--no-spatial-order]\n"
        "  sim                 run the Monte-Carlo simulation of each config in turn\n"
        "  -c, --configs       config files (INI), as many as needed\n"
        "  -d, --device        GPU device id (default 0)\n"
        "  -p, --use_cpu       not supported by this engine\n"
        "  -g, --gpu_info      print GPU information\n"
        "  -v, --version       print the version\n");
}

int main(int argc, char **argv) {
    if (argc < 2) { usage(); return 0; }
    std::vector<std::string> cfgs;
    bool sim = false, use_cpu = false;
    uint32_t dev = 0;
    int32_t rng = SW_RNG_PHILOX, prec = SW_PREC_F32, field = SW_FIELD_F32, order = 1;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto need = [&](const char *what) -> const char * {
            if (i + 1 >= argc) { std::fprintf(stderr, "%s needs a value\n", what); std::exit(2); }
            return argv[++i];
        };
        if (a == "-v" || a == "--version") { std::printf("%s\n", sw_version()); return 0; }
        if (a == "-g" || a == "--gpu_info") {
            const int n = sw_device_count();
            std::printf("Number of GPU(s): %d\n", n);
            for (int d = 0; d < n; d++) {
                char buf[1024];
                if (sw_device_info(d, buf, sizeof buf) == SW_OK) std::printf("%s", buf);
            }
            return 0;
        }
        if (a == "-h" || a == "--help") { usage(); return 0; }
        if (a == "sim") { sim = true; continue; }
        if (a == "-c" || a == "--configs") {
            while (i + 1 < argc && argv[i + 1][0] != '-') cfgs.push_back(argv[++i]);
            continue;
        }
        if (a == "-p" || a == "--use_cpu") { use_cpu = true; continue; }
        if (a == "-d" || a == "--device") { dev = (uint32_t)std::atoi(need("-d")); continue; }
        if (a == "--rng") { const std::string v = need("--rng"); rng = v == "replay" ? SW_RNG_REPLAY : SW_RNG_PHILOX; continue; }
        if (a == "--precision") { const std::string v = need("--precision"); prec = v == "kahan" ? SW_PREC_F32_KAHAN : SW_PREC_F32; continue; }
        if (a == "--field") {
            const std::string v = need("--field");
            field = v == "f16" ? SW_FIELD_F16 : (v == "bf16" ? SW_FIELD_BF16 : SW_FIELD_F32);
            continue;
        }
        if (a == "--no-spatial-order") { order = 0; continue; }
        std::fprintf(stderr, "unknown argument: %s\n", a.c_str());
        usage();
        return 2;
    }
    if (!sim) { usage(); return 0; }
    if (cfgs.empty()) { std::fprintf(stderr, "sim needs -c CONFIG\n"); return 2; }
    std::printf("SpinWalk (MI355X engine) %s\n", sw_version());
    if (use_cpu) {
        std::fprintf(stderr, "Error: this engine has no CPU path (-p)\n");
        return 1;
    }
    sw_engine *eng = nullptr;
    if (sw_engine_create((int)dev, &eng) != SW_OK) {
        std::fprintf(stderr, "Error: %s\n", sw_last_error(nullptr));
        return 1;
    }
    int rc = SW_OK;
    for (auto &c : cfgs) {
        std::printf("<%s>\n", c.c_str());
        rc = sw_run_config_ex(eng, c.c_str(), rng, prec, field, order);
        if (rc != SW_OK) {
            std::fprintf(stderr, "Error: Simulation failed: %s\n", sw_host_last_error());
            break;
        }
    }
    sw_engine_destroy(eng);
    if (rc == SW_OK) std::printf("Simulation completed successfully.\n");
    return rc == SW_OK ? 0 : 1;
}
