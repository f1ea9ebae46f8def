// bh_phase.cpp -- command-line driver with the reference's options (src/main.cpp:100-177).
//
//   bhx-phase -m SITES -n BOSONS -J HOPPING -U INTERACTION -u POTENTIAL -r RANGE -s STEP -f J|U|u -t exact
//   bhx-phase -t mean -i STEPS -e PRECISION
//
// -t exact runs Analysis::exact_parameters and writes phase.txt; -t mean runs
// Analysis::mean_field_parameters (an STEPS x STEPS grid, convergence threshold 10^-PRECISION) and writes
// mean_field.txt (src/main.cpp:179-201).  Validation follows the reference: the calculation type must be
// exact or mean, s < r, and the fixed parameter J, U or u, with exit code 1 otherwise; -i and -e are read
// as integers (the reference stores them in ints) and default to 100 and 6 (the reference leaves them
// uninitialised).  Extensions: --out FILE, --prec fp64|mixed, --reference-conventions (the reference's
// doubled hopping and sum n(n+1) interaction), --seed N (start values of the mean-field grid).
#include <getopt.h>

#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "bh_host.hpp"

static void usage() {
    std::cout << "Usage: bhx-phase [options]\n"
                 "  -m, --sites N            number of lattice sites (periodic chain)\n"
                 "  -n, --bosons N           number of bosons\n"
                 "  -J, --hopping X          hopping parameter\n"
                 "  -U, --interaction X      on-site interaction\n"
                 "  -u, --potential X        chemical potential\n"
                 "  -r, --range X            range of the varying parameters\n"
                 "  -s, --step X             step of the varying parameters (s < r)\n"
                 "  -f, --fixed J|U|u        parameter held fixed\n"
                 "  -t, --type exact|mean    calculation type\n"
                 "  -i, --iterations N       mean field: N x N grid of (mu, J) (default 100)\n"
                 "  -e, --epsilon P          mean field: convergence threshold 10^-P (default 6)\n"
                 "      --out FILE           output file (default phase.txt, mean_field.txt)\n"
                 "      --seed N             mean field: seed of the start values (default 0)\n"
                 "      --prec fp64|mixed    eigensolver precision (default mixed)\n"
                 "      --reference-conventions  doubled hopping and sum n(n+1) interaction\n"
                 "  -h, --help               this text\n";
}

int main(int argc, char* argv[]) {
    int m = 0, n = 0;
    double J = 0, U = 0, mu = 0, s = 0, r = 0;
    int iters = 100, precision = 6;
    unsigned long long seed = 0;
    std::string fixed_param, calc_type, out;
    enum { OPT_OUT = 1000, OPT_PREC, OPT_REFCONV, OPT_SEED };
    const char* const short_opts = "m:n:J:U:u:r:s:f:t:i:e:h";
    const option long_opts[] = {{"sites", required_argument, nullptr, 'm'},
                                {"bosons", required_argument, nullptr, 'n'},
                                {"hopping", required_argument, nullptr, 'J'},
                                {"interaction", required_argument, nullptr, 'U'},
                                {"potential", required_argument, nullptr, 'u'},
                                {"range", required_argument, nullptr, 'r'},
                                {"step", required_argument, nullptr, 's'},
                                {"fixed", required_argument, nullptr, 'f'},
                                {"type", required_argument, nullptr, 't'},
                                {"iterations", required_argument, nullptr, 'i'},
                                {"epsilon", required_argument, nullptr, 'e'},
                                {"out", required_argument, nullptr, OPT_OUT},
                                {"prec", required_argument, nullptr, OPT_PREC},
                                {"reference-conventions", no_argument, nullptr, OPT_REFCONV},
                                {"seed", required_argument, nullptr, OPT_SEED},
                                {"help", no_argument, nullptr, 'h'},
                                {nullptr, no_argument, nullptr, 0}};
    try {
        for (;;) {
            const int opt = getopt_long(argc, argv, short_opts, long_opts, nullptr);
            if (opt == -1) break;
            switch (opt) {
                case 'm': m = std::stoi(optarg); break;
                case 'n': n = std::stoi(optarg); break;
                case 'J': J = std::stod(optarg); break;
                case 'U': U = std::stod(optarg); break;
                case 'u': mu = std::stod(optarg); break;
                case 'r': r = std::stod(optarg); break;
                case 's': s = std::stod(optarg); break;
                case 'f': fixed_param = optarg; break;
                case 't': calc_type = optarg; break;
                case 'i': iters = static_cast<int>(std::stod(optarg)); break;
                case 'e': precision = static_cast<int>(std::stod(optarg)); break;
                case OPT_SEED: seed = std::stoull(optarg); break;
                case OPT_OUT: out = optarg; break;
                case OPT_PREC:
                    if (std::strcmp(optarg, "fp64") == 0) bhx::settings().prec = BH_PREC_FP64;
                    else if (std::strcmp(optarg, "mixed") == 0) bhx::settings().prec = BH_PREC_MIXED;
                    else {
                        std::cerr << "Error: --prec must be fp64 or mixed." << std::endl;
                        return 1;
                    }
                    break;
                case OPT_REFCONV: bhx::settings().flags = BH_CONV_REF_HOP | BH_CONV_REF_INTERACTION; break;
                case 'h':
                default: usage(); return 0;
            }
        }
    } catch (const std::exception&) {
        std::cerr << "Error: invalid numeric argument." << std::endl;
        return 1;
    }
    if (calc_type != "exact" && calc_type != "mean") {
        std::cerr << "Error: calculation type must be 'exact' or 'mean'." << std::endl;
        return 1;
    }
    if (calc_type == "mean") {
        if (iters < 1) {
            std::cerr << "Error: the number of iterations must be positive." << std::endl;
            return 1;
        }
        if (out.empty()) out = "mean_field.txt";
        try {
            const int64_t bad = Analysis::mean_field_parameters(iters, precision, out, seed);
            if (bad > 0)
                std::cerr << "Warning: " << bad << " points did not converge within the iteration caps." << std::endl;
        } catch (const std::exception& e) {
            std::cerr << "Error: " << e.what() << std::endl;
            return 3;
        }
        std::cout << "wrote " << out << std::endl;
        return 0;
    }
    if (out.empty()) out = "phase.txt";
    if (s >= r) {
        std::cerr << "Error: s must be smaller than r." << std::endl;
        return 1;
    }
    if (fixed_param != "J" && fixed_param != "U" && fixed_param != "u") {
        std::cerr << "Error: fixed parameter must be J, U or u." << std::endl;
        return 1;
    }
    if (m < 1 || n < 0 || s <= 0) {
        std::cerr << "Error: need m >= 1, n >= 0 and s > 0." << std::endl;
        return 1;
    }
    try {
        Analysis::exact_parameters(m, n, J, U, mu, s, r, fixed_param, out);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 3;
    }
    std::cout << "wrote " << out << std::endl;
    return 0;
}
