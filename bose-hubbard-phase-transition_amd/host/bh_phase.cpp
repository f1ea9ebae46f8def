// bh_phase.cpp -- command-line driver with the reference's options (src/main.cpp:100-177).
//
//   bhx-phase -m SITES -n BOSONS -J HOPPING -U INTERACTION -u POTENTIAL -r RANGE -s STEP -f J|U|u -t exact
//
// runs Analysis::exact_parameters and writes phase.txt.  Validation follows the reference: the
// calculation type must be exact or mean, s < r, and the fixed parameter J, U or u, with exit code 1
// otherwise.  -t mean (the mean-field path) is not part of this build and exits with code 2.  -i and -e
// are accepted and ignored, as the exact path of the reference ignores them.  Extensions: --out FILE,
// --prec fp64|mixed, --reference-conventions (the reference's doubled hopping and sum n(n+1) interaction).
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
                 "  -t, --type exact|mean    calculation type (exact only in this build)\n"
                 "  -i, --iterations N       accepted, unused by the exact path\n"
                 "  -e, --epsilon X          accepted, unused by the exact path\n"
                 "      --out FILE           output file (default phase.txt)\n"
                 "      --prec fp64|mixed    eigensolver precision (default mixed)\n"
                 "      --reference-conventions  doubled hopping and sum n(n+1) interaction\n"
                 "  -h, --help               this text\n";
}

int main(int argc, char* argv[]) {
    int m = 0, n = 0;
    double J = 0, U = 0, mu = 0, s = 0, r = 0;
    std::string fixed_param, calc_type, out = "phase.txt";
    enum { OPT_OUT = 1000, OPT_PREC, OPT_REFCONV };
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
                case 'i':
                case 'e': break;
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
        std::cerr << "Error: the mean-field calculation is not part of this build." << std::endl;
        return 2;
    }
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
