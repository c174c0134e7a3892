// driver_common.hpp -- command line and file formats shared by the drivers
// that restate the reference's programs over HipImpl<T>.
//
// Options: the Boost.program_options set of include/Command_Parser.hpp:21-36
// (same names, short forms, defaults and help text), parsed here without
// Boost: "-m 64", "-m64", "--M 64", "--M=64".  Added: --init FILE (initial
// guess) and --seed S (seed of the uniform initial guess).
// Files: raw float32, include/detail/CPRA.inl:8-48.
#pragma once
#include <cpra_hip.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace cpra_driver {

struct Options {
    int M = 160, N = 160, L = 160, P = 90, Epi = 1, Iter = 1, PreIter = 1000, Batch = 1, Total = 100;
    float Beta = 0.9f;
    std::string data_constraint = "./data_constraint.bin";
    std::string space_constraint = "./space_constraint.bin";
    std::string output_reconstruction = "./output_reconstruction.bin";
    std::string output_error = "./output_error.bin";
    std::string init;
    unsigned long long seed = 0;
};

inline const char* kHelp =
    "Allowed options:\n"
    "  -h [ --help ]                         produce help message\n"
    "  -m [ --M ] arg (=160)                 set first dimension\n"
    "  -n [ --N ] arg (=160)                 set second dimension\n"
    "  -l [ --L ] arg (=160)                 set third dimension\n"
    "  -p [ --P ] arg (=90)                  set number of projected objects\n"
    "  -e [ --epi ] arg (=1)                 set number of episodes\n"
    "  -i [ --iter ] arg (=1)                set number of iterations within each episodes\n"
    "  --preit arg (=1000)                   set number of iterations for pre-reconstruction\n"
    "  --Batch arg (=1)                      set batch size\n"
    "  --Total arg (=100)                    Total batch size\n"
    "  --Beta arg (=0.9)                     set shrinkwrap beta value\n"
    "  -d [ --data_constraint ] arg (=./data_constraint.bin)\n"
    "                                        Data Constraint file\n"
    "  -s [ --space_constraint ] arg (=./space_constraint.bin)\n"
    "                                        Space Constraint file\n"
    "  -r [ --output_reconstruction ] arg (=./output_reconstruction.bin)\n"
    "                                        Output file for reconstructions\n"
    "  --output_error arg (=./output_error.bin)\n"
    "                                        Output file for errors\n"
    "  --init arg                            initial guess file (complex64)\n"
    "  --seed arg (=0)                       seed of the uniform initial guess\n";

// name used in messages; set by each driver's main
inline const char*& program_name()
{
    static const char* name = "cpra";
    return name;
}

[[noreturn]] inline void usage_error(const std::string& msg)
{
    std::cerr << program_name() << ": " << msg << "\n" << kHelp;
    std::exit(2);
}

inline int to_int(const std::string& v, const std::string& name)
{
    char* end = nullptr;
    const long x = std::strtol(v.c_str(), &end, 10);
    if (v.empty() || *end) usage_error("option " + name + " needs an integer, got '" + v + "'");
    return (int)x;
}

// Boost-style options: "-m 64", "-m64", "--M 64", "--M=64".
// min_P: the smallest number of projections the driver accepts (the carousel
// warm start reads both neighbours of a projection, so it needs 2).
inline Options parse(int argc, char** argv, int min_P = 1)
{
    Options o;
    struct Spec {
        const char* shrt;
        const char* lng;
        int kind;  // 0 int, 1 float, 2 string, 3 u64
        void* dst;
    };
    const Spec specs[] = {
        {"m", "M", 0, &o.M},
        {"n", "N", 0, &o.N},
        {"l", "L", 0, &o.L},
        {"p", "P", 0, &o.P},
        {"e", "epi", 0, &o.Epi},
        {"i", "iter", 0, &o.Iter},
        {nullptr, "preit", 0, &o.PreIter},
        {nullptr, "Batch", 0, &o.Batch},
        {nullptr, "Total", 0, &o.Total},
        {nullptr, "Beta", 1, &o.Beta},
        {"d", "data_constraint", 2, &o.data_constraint},
        {"s", "space_constraint", 2, &o.space_constraint},
        {"r", "output_reconstruction", 2, &o.output_reconstruction},
        {nullptr, "output_error", 2, &o.output_error},
        {nullptr, "init", 2, &o.init},
        {nullptr, "seed", 3, &o.seed},
    };
    for (int k = 1; k < argc; ++k) {
        std::string a = argv[k];
        if (a == "-h" || a == "--help") {
            std::cout << kHelp;
            std::exit(EXIT_SUCCESS);
        }
        std::string name, value;
        bool have_value = false;
        const Spec* sp = nullptr;
        if (a.rfind("--", 0) == 0) {
            name = a.substr(2);
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
                have_value = true;
            }
            for (const Spec& s : specs)
                if (name == s.lng) sp = &s;
        } else if (a.size() >= 2 && a[0] == '-') {
            name = a.substr(1, 1);
            if (a.size() > 2) {
                value = a.substr(2);
                have_value = true;
            }
            for (const Spec& s : specs)
                if (s.shrt && name == s.shrt) sp = &s;
        }
        if (!sp) usage_error("unrecognised option '" + a + "'");
        if (!have_value) {
            if (k + 1 >= argc) usage_error("option '" + a + "' needs a value");
            value = argv[++k];
        }
        switch (sp->kind) {
        case 0: *static_cast<int*>(sp->dst) = to_int(value, name); break;
        case 1: *static_cast<float*>(sp->dst) = std::strtof(value.c_str(), nullptr); break;
        case 2: *static_cast<std::string*>(sp->dst) = value; break;
        case 3: *static_cast<unsigned long long*>(sp->dst) = std::strtoull(value.c_str(), nullptr, 10); break;
        }
    }
    if (o.M < 1 || o.N < 1 || o.L < 1 || o.P < min_P || o.Epi < 0 || o.Iter < 0 || o.PreIter < 0 ||
        o.Batch < 1 || o.Total < 0)
        usage_error("sizes must be positive, P >= " + std::to_string(min_P) +
                    ", iteration counts >= 0");
    if (o.Beta == 0.f) usage_error("Beta must not be 0");
    return o;
}

// include/detail/CPRA.inl:8-29 (raw float32; a short file is an error here)
inline std::vector<float> read_floats(const std::string& file, size_t count)
{
    std::ifstream f(file, std::ios::binary);
    if (!f) throw std::runtime_error("Reading input file " + file + " not exist!");
    std::vector<float> v(count);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(count * sizeof(float)));
    if ((size_t)f.gcount() != count * sizeof(float))
        throw std::runtime_error("input file " + file + " holds fewer than " + std::to_string(count) + " floats");
    return v;
}

// include/detail/CPRA.inl:31-48
inline void write_floats(const std::string& file, const float* p, size_t count)
{
    if (std::ifstream(file).good()) std::cout << "We're over writting file " << file << "!" << std::endl;
    std::ofstream f(file, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), (std::streamsize)(count * sizeof(float)));
    if (!f) throw std::runtime_error("writing " + file + " failed");
}

inline void check(int rc, cpra_hip_handle h, const char* what)
{
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + cpra_hip_last_error(h));
}

}  // namespace cpra_driver
