// Executable counterpart of the reference's generated main
// (kernel_utils/common/kernel.cpp.in:645-748):
//
//     pyjac_hip_main <build_dir> <problem_size> <work_size> [compile]
//
// reads <build_dir>/pyjac_hip_main.cfg (written by `pyjac --lang hip`):
//
//     tables <file>        the mechanism tables (MechTables.save_binary)
//     kernel jacobian | species_rates
//     conp 0|1   order C|F   jac_format 0|1   jac_type 0|1|2
//     data <file>          initial conditions (read_initial_conditions)
//     for_validation 0|1   write the outputs as <build_dir>/<name>.bin
//
// then reads problem_size states, evaluates the kernel through the
// host-pointer entry points in chunks of work_size states (0: all at once),
// and prints "problem_size,compilation_time,setup_time,runtime" (ms; the
// compilation time is -1: nothing is compiled at run time).  Errors print the
// library's message and return 1 instead of the reference's exit(-1).

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/pyjac_hip.h"

namespace {

std::map<std::string, std::string> read_config(const std::string& path) {
    std::map<std::string, std::string> kv;
    FILE* fp = std::fopen(path.c_str(), "r");
    if (!fp) return kv;
    char key[256], val[4096];
    while (std::fscanf(fp, "%255s %4095s", key, val) == 2) kv[key] = val;
    std::fclose(fp);
    return kv;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int report(const char* what, int rc) {
    std::fprintf(stderr, "pyjac_hip_main: %s failed (status %d): %s\n", what, rc,
                 pyjac_hip_last_error());
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <build_dir> <problem_size> <work_size> [compile]\n",
                     argv[0]);
        return 1;
    }
    const std::string dir = argv[1];
    const size_t problem_size = std::strtoull(argv[2], nullptr, 10);
    const size_t work_size = std::strtoull(argv[3], nullptr, 10);
    auto cfg = read_config(dir + "/pyjac_hip_main.cfg");
    if (cfg.empty()) {
        std::fprintf(stderr, "pyjac_hip_main: no %s/pyjac_hip_main.cfg\n", dir.c_str());
        return 1;
    }
    auto path = [&](const std::string& key) {
        const std::string& v = cfg[key];
        return (v.empty() || v[0] == '/') ? v : dir + "/" + v;
    };
    const bool jac = cfg["kernel"] == "jacobian";
    const int conp = std::atoi(cfg["conp"].c_str());
    const char order = cfg["order"].empty() ? 'C' : cfg["order"][0];
    const int jac_format = std::atoi(cfg["jac_format"].c_str());
    const int jac_type = std::atoi(cfg["jac_type"].c_str());
    const bool validate = std::atoi(cfg["for_validation"].c_str()) != 0;

    pyjac_mech_tables* tables = nullptr;
    int rc = pyjac_hip_tables_load(path("tables").c_str(), &tables);
    if (rc) return report("loading the tables", rc);
    const int NN = tables->n_species + 1;
    std::vector<double> param(problem_size), phi(problem_size * NN);
    rc = pyjac_hip_read_initial_conditions(path("data").c_str(), problem_size, NN, param.data(),
                                           phi.data(), order);
    if (rc) return report("reading the initial conditions", rc);

    auto t0 = std::chrono::steady_clock::now();
    pyjac_handle* h = nullptr;
    rc = pyjac_hip_create(tables, conp, order, jac_format, jac_type, &h);
    if (!rc) rc = pyjac_hip_resize(h, problem_size, work_size);
    if (rc) return report("setting up the kernel", rc);
    const double setup_time = ms_since(t0);

    const size_t width = jac ? (size_t)pyjac_hip_jacobian_size(h) : (size_t)NN;
    std::vector<double> out(problem_size * width);
    t0 = std::chrono::steady_clock::now();
    rc = jac ? pyjac_hip_jacobian(h, param.data(), phi.data(), out.data())
             : pyjac_hip_species_rates(h, param.data(), phi.data(), out.data());
    if (rc) return report("the kernel call", rc);
    const double runtime = ms_since(t0);

    std::printf("%zu,%.15le,%.15le,%.15le\n", problem_size, -1.0, setup_time, runtime);
    if (validate) {
        const std::string name = dir + (jac ? "/jac.bin" : "/dphi.bin");
        rc = pyjac_hip_write_data(name.c_str(), out.data(), out.size());
        if (rc) return report("writing the output", rc);
    }
    pyjac_hip_destroy(h);
    pyjac_hip_tables_free(tables);
    return 0;
}
