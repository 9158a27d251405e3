// spinwalk_cli.cpp -- `spinwalk sim -c config.ini ...` on the MI355X engine, and
// the host-only `config` and `dwi` sequence generators.
//
// The command-line surface of the reference's sim subcommand and global flags
// (src/spinwalk.cpp:45-56): -c/--configs (one or more files), -p/--use_cpu,
// -d/--device, -g/--gpu_info, -v/--version.  -p is refused: this engine has no CPU
// path.  Extra sim options select the engine's modes: --rng philox|replay,
// --precision f32|kahan, --field f32|f16|bf16|q16, --no-spatial-order.
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
        "                [--precision f32|kahan] [--field f32|f16|bf16|q16] [--no-spatial-order]\n"
        "       spinwalk config -s GRE|SE|bSSFP -p PHANTOM... -e TE_us -t TIMESTEP_us -o OUTPUT\n"
        "       spinwalk dwi -b B... -v X Y Z -d START_ms delta_ms DELTA_ms -c CONFIG\n"
        "  sim                 run the Monte-Carlo simulation of each config in turn\n"
        "  config              write a GRE, SE or bSSFP config and its default_config.ini parent\n"
        "  dwi                 insert a PGSE gradient table (b-values in s/mm^2) into a config\n"
        "  -c, --configs       config files (INI), as many as needed\n"
        "  -d, --device        GPU device id (default 0)\n"
        "  -p, --use_cpu       not supported by this engine\n"
        "  -g, --gpu_info      print GPU information\n"
        "  -v, --version       print the version\n");
}

static bool is_number(const char *a) {
    char *end = nullptr;
    std::strtod(a, &end);
    return end && end != a && *end == 0;
}

// `spinwalk config ...` (src/spinwalk.cpp:74-79, src/config/handler.cpp:13-39)
static int run_config_generator(int argc, char **argv) {
    std::string seq, out;
    std::vector<std::string> ph;
    long te = -1, dt = -1;
    for (int i = 2; i < argc; i++) {
        const std::string a = argv[i];
        const bool more = i + 1 < argc;
        if ((a == "-s" || a == "--seq_name") && more) seq = argv[++i];
        else if ((a == "-e" || a == "--TE") && more) te = std::atol(argv[++i]);
        else if ((a == "-t" || a == "--timestep") && more) dt = std::atol(argv[++i]);
        else if ((a == "-o" || a == "--output") && more) out = argv[++i];
        else if (a == "-p" || a == "--phantoms") { while (i + 1 < argc && argv[i + 1][0] != '-') ph.push_back(argv[++i]); }
        else { std::fprintf(stderr, "config: unknown or incomplete argument %s\n", a.c_str()); return 2; }
    }
    if (seq.empty() || out.empty() || ph.empty() || te <= 0 || dt <= 0) {
        std::fprintf(stderr, "config needs -s, -p, -e (positive), -t (positive) and -o\n");
        return 2;
    }
    std::vector<const char *> pp;
    for (auto &x : ph) pp.push_back(x.c_str());
    if (sw_generate_config(seq.c_str(), (uint32_t)te, (uint32_t)dt, pp.data(), (uint32_t)pp.size(), out.c_str()) != SW_OK) {
        std::fprintf(stderr, "Error: Configuration file generation failed: %s\n", sw_host_last_error());
        return 1;
    }
    std::printf("Configuration file is generated in %s\n", out.c_str());
    return 0;
}

// `spinwalk dwi ...` (src/spinwalk.cpp:81-85, src/dwi/handler.cpp:6-17)
static int run_dwi_generator(int argc, char **argv) {
    std::vector<double> b;
    std::vector<float> dir;
    std::vector<long> delta;
    std::string cfg;
    for (int i = 2; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "-b" || a == "--bvalue") { while (i + 1 < argc && is_number(argv[i + 1])) b.push_back(std::atof(argv[++i])); }
        else if (a == "-v" || a == "--bvector") { while (i + 1 < argc && is_number(argv[i + 1])) dir.push_back((float)std::atof(argv[++i])); }
        else if (a == "-d" || a == "--delta") { while (i + 1 < argc && is_number(argv[i + 1])) delta.push_back(std::atol(argv[++i])); }
        else if ((a == "-c" || a == "--config") && i + 1 < argc) cfg = argv[++i];
        else { std::fprintf(stderr, "dwi: unknown or incomplete argument %s\n", a.c_str()); return 2; }
    }
    if (b.empty() || dir.size() != 3 || delta.size() != 3 || cfg.empty()) {
        std::fprintf(stderr, "dwi needs -b B..., -v X Y Z, -d START delta DELTA and -c CONFIG\n");
        return 2;
    }
    for (long v : delta)
        if (v < 0) { std::fprintf(stderr, "dwi: times must not be negative\n"); return 2; }
    std::printf("Generating PGSE gradient table...\n");
    if (sw_generate_pgse(cfg.c_str(), b.data(), (uint32_t)b.size(), dir.data(), (uint32_t)delta[0], (uint32_t)delta[1],
                         (uint32_t)delta[2]) != SW_OK) {
        std::fprintf(stderr, "Error: Diffusion gradient generation failed: %s\n", sw_host_last_error());
        return 1;
    }
    std::printf("Diffusion gradient table is generated successfully.\n");
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) { usage(); return 0; }
    if (std::strcmp(argv[1], "config") == 0) return run_config_generator(argc, argv);
    if (std::strcmp(argv[1], "dwi") == 0) return run_dwi_generator(argc, argv);
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
        if (a == "--rng") {
            const std::string v = need("--rng");
            if (v != "philox" && v != "replay") { std::fprintf(stderr, "unknown rng: %s\n", v.c_str()); return 2; }
            rng = v == "replay" ? SW_RNG_REPLAY : SW_RNG_PHILOX;
            continue;
        }
        if (a == "--precision") {
            const std::string v = need("--precision");
            if (v != "f32" && v != "kahan") { std::fprintf(stderr, "unknown precision: %s\n", v.c_str()); return 2; }
            prec = v == "kahan" ? SW_PREC_F32_KAHAN : SW_PREC_F32;
            continue;
        }
        if (a == "--field") {
            const std::string v = need("--field");
            if (v == "f32") field = SW_FIELD_F32;
            else if (v == "f16") field = SW_FIELD_F16;
            else if (v == "bf16") field = SW_FIELD_BF16;
            else if (v == "q16") field = SW_FIELD_Q16;
            else { std::fprintf(stderr, "unknown field format: %s\n", v.c_str()); return 2; }
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
