// opts.cpp -- TiltRec's getopt_long interface (src/TiltRec/src/opts/opts.cpp:
// 126-360): the same long and short options, the same comma-separated
// --geometry and --method values, and the misspelled --inital (kept, with
// --initial as an alias).  Where the reference reads uninitialised values
// (a bare "--method SIRT", opts.cpp:112-124) or accepts nonsense (an axis
// check that is always false, :267), this parser reports an error.
#include "opts.h"

#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <vector>

namespace trec_cli {

namespace {

std::vector<std::string> split(const std::string &s, char sep)
{
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string t;
    while (std::getline(ss, t, sep)) out.push_back(t);
    if (!s.empty() && s.back() == sep) out.push_back("");
    return out;
}

bool to_float(const std::string &s, float &v)
{
    char *end = nullptr;
    v = strtof(s.c_str(), &end);
    return !s.empty() && end && *end == '\0';
}

bool to_int(const std::string &s, int &v)
{
    char *end = nullptr;
    long r = strtol(s.c_str(), &end, 10);
    v = (int)r;
    return !s.empty() && end && *end == '\0';
}

// one whitespace-free token, as `iss >> opts_->input` accepts (opts.cpp:160-174)
bool one_token(const char *arg, std::string &out, const char *what)
{
    std::stringstream iss(arg);
    iss >> out;
    std::string extra;
    if (iss.fail() || out.empty()) {
        fprintf(stderr, "Invalid argument for %s '%s'.\n", what, arg);
        return false;
    }
    if (iss >> extra) {
        fprintf(stderr, "Too many arguments for %s '%s'.\n", what, arg);
        return false;
    }
    return true;
}

std::string esc(const std::string &s)
{
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

}  // namespace

void usage()
{
    std::cout << "[--input(-i) Input Filename]\n    MRC file for reconstruction\n"
                 "[--output(-o) Output Filename]\n    MRC filename for result\n"
                 "[--tiltfile(-t) Tilt Angle Filename]\n    Tilt Angles\n"
                 "[--inital(-n) / --initial Initial Reconstruction Filename]\n"
                 "    MRC file as initial model for the reconstruction (optional)\n"
                 "[--geometry(-g) offset,pitch_angle,zshift,thickness]\n"
                 "[--method(-m) Method]\n"
                 "    BPT | FBP | WBP | SIRT,iterations,relax | SART,iterations,relax |\n"
                 "    ADMM,iterations,cgiter,relax,soft\n"
                 "[--axis(-a) y|z]  volume orientation: y = TiltRec (X,thickness,Y), z = TiltRecZ (X,Y,thickness)\n"
                 "[--f2b(-f)]  byte output: (v - mean) / sd clamped to +-10, mapped to 0..255\n"
                 "[--rules mpi|cuda]  thresholds and y origin of TiltRec-mpi or TiltRec-cuda\n"
                 "[--storage f32|f16|bf16]  projection/residual storage of SIRT and SART\n"
                 "[--device N]  HIP device\n"
                 "[--output-mode 2|12]  MRC mode of the volume: 2 float (default), 12 half\n"
                 "[--mrc2014]  MRC2014 header: MAP and machine stamps, sampling, whole-map statistics\n"
                 "-help(-h)\n"
                 "EXAMPLE:\n"
                 "  TiltRec-cuda --input BBb_fin.mrc --output BBb_SIRT.mrc --tiltfile BBb.rawtlt "
                 "--geometry 0,0,0,300 --method SIRT,20,1.0\n";
}

int get_opts(int argc, char **argv, Options &o)
{
    enum { OPT_RULES = 1000, OPT_STORAGE, OPT_DEVICE, OPT_DRY, OPT_OUTMODE, OPT_MRC2014 };
    static struct option longopts[] = {{"help", no_argument, nullptr, 'h'},
                                       {"input", required_argument, nullptr, 'i'},
                                       {"output", required_argument, nullptr, 'o'},
                                       {"inital", required_argument, nullptr, 'n'},
                                       {"initial", required_argument, nullptr, 'n'},
                                       {"tiltfile", required_argument, nullptr, 't'},
                                       {"geometry", required_argument, nullptr, 'g'},
                                       {"method", required_argument, nullptr, 'm'},
                                       {"axis", required_argument, nullptr, 'a'},
                                       {"f2b", no_argument, nullptr, 'f'},
                                       {"rules", required_argument, nullptr, OPT_RULES},
                                       {"storage", required_argument, nullptr, OPT_STORAGE},
                                       {"device", required_argument, nullptr, OPT_DEVICE},
                                       {"dry-run", no_argument, nullptr, OPT_DRY},
                                       {"output-mode", required_argument, nullptr, OPT_OUTMODE},
                                       {"mrc2014", no_argument, nullptr, OPT_MRC2014},
                                       {nullptr, 0, nullptr, 0}};
    optind = 1;
    opterr = 0;
    int ch;
    while ((ch = getopt_long(argc, argv, "hi:o:n:t:g:m:a:f", longopts, nullptr)) != -1) {
        switch (ch) {
        case 'h': usage(); return 0;
        case 'i':
            if (!one_token(optarg, o.input, "input")) return -1;
            break;
        case 'o':
            if (!one_token(optarg, o.output, "output")) return -1;
            break;
        case 'n':
            if (!one_token(optarg, o.initial, "initial")) return -1;
            break;
        case 't':
            if (!one_token(optarg, o.angle, "tiltfile")) return -1;
            break;
        case 'g': {
            auto p = split(optarg, ',');
            if (p.size() != 4 || !to_float(p[0], o.offset) || !to_float(p[1], o.pitch) || !to_float(p[2], o.zshift) ||
                !to_int(p[3], o.thickness)) {
                fprintf(stderr, "Invalid argument for geometry '%s' (offset,pitch,zshift,thickness).\n", optarg);
                return -1;
            }
            break;
        }
        case 'm': {
            auto p = split(optarg, ',');
            o.method = p[0];
            if (o.method == "SIRT" || o.method == "SART") {
                if (p.size() != 3 || !to_int(p[1], o.iteration) || !to_float(p[2], o.gamma)) {
                    fprintf(stderr, "Iteration Method: [--method/-m %s,iterations,relax]\n", o.method.c_str());
                    return -1;
                }
            } else if (o.method == "ADMM") {
                if (p.size() != 5 || !to_int(p[1], o.iteration) || !to_int(p[2], o.cgiter) ||
                    !to_float(p[3], o.gamma) || !to_float(p[4], o.soft)) {
                    fprintf(stderr, "ADMM Method: [--method/-m ADMM,iterations,cgiter,relax,soft]\n");
                    return -1;
                }
            } else if (o.method == "BPT" || o.method == "FBP" || o.method == "WBP") {
                if (p.size() != 1) {
                    fprintf(stderr, "Too many arguments for method '%s'.\n", optarg);
                    return -1;
                }
            } else {
                fprintf(stderr, "Invalid argument for method '%s'.\n", optarg);
                return -1;
            }
            break;
        }
        case 'a':
            if (!one_token(optarg, o.axis, "axis")) return -1;
            if (o.axis != "y" && o.axis != "Y" && o.axis != "z" && o.axis != "Z") {
                fprintf(stderr, "Invalid argument for axis '%s'.\n", optarg);
                return -1;
            }
            break;
        case 'f': o.f2b = true; break;
        case OPT_RULES:
            o.rules = optarg;
            if (o.rules != "mpi" && o.rules != "cuda") {
                fprintf(stderr, "Invalid argument for rules '%s' (mpi or cuda).\n", optarg);
                return -1;
            }
            break;
        case OPT_STORAGE:
            o.storage = optarg;
            if (o.storage != "f32" && o.storage != "f16" && o.storage != "bf16") {
                fprintf(stderr, "Invalid argument for storage '%s' (f32, f16 or bf16).\n", optarg);
                return -1;
            }
            break;
        case OPT_DEVICE:
            if (!to_int(optarg, o.device) || o.device < 0) {
                fprintf(stderr, "Invalid argument for device '%s'.\n", optarg);
                return -1;
            }
            break;
        case OPT_DRY: o.dry_run = true; break;
        case OPT_OUTMODE:
            if (!to_int(optarg, o.output_mode) || (o.output_mode != 2 && o.output_mode != 12)) {
                fprintf(stderr, "Invalid argument for output-mode '%s' (2 or 12).\n", optarg);
                return -1;
            }
            break;
        case OPT_MRC2014: o.mrc2014 = true; break;
        default:
            fprintf(stderr, "Invalid option '%s'.\n", argv[optind - 1]);
            return -1;
        }
    }
    if (optind < argc) {
        fprintf(stderr, "Unexpected argument '%s'.\n", argv[optind]);
        return -1;
    }
    if (o.f2b && o.output_mode != 2) {
        fprintf(stderr, "--f2b writes bytes; it cannot be combined with --output-mode %d.\n", o.output_mode);
        return -1;
    }
    if (o.input.empty() || o.output.empty() || o.angle.empty() || o.method.empty() || o.thickness <= 0) {
        fprintf(stderr, "--input, --output, --tiltfile, --geometry (thickness > 0) and --method are required.\n");
        return -1;
    }
    return 1;
}

std::string opts_json(const Options &o)
{
    std::ostringstream s;
    s << "{\"input\":\"" << esc(o.input) << "\",\"output\":\"" << esc(o.output) << "\",\"initial\":\""
      << esc(o.initial) << "\",\"tiltfile\":\"" << esc(o.angle) << "\",\"offset\":" << o.offset
      << ",\"pitch\":" << o.pitch << ",\"zshift\":" << o.zshift << ",\"thickness\":" << o.thickness
      << ",\"method\":\"" << esc(o.method) << "\",\"iterations\":" << o.iteration << ",\"gamma\":" << o.gamma
      << ",\"cgiter\":" << o.cgiter << ",\"soft\":" << o.soft
      << ",\"axis\":\"" << esc(o.axis) << "\",\"f2b\":" << (o.f2b ? "true" : "false") << ",\"rules\":\""
      << esc(o.rules) << "\",\"storage\":\"" << esc(o.storage) << "\",\"device\":" << o.device
      << ",\"output_mode\":" << o.output_mode << ",\"mrc2014\":" << (o.mrc2014 ? "true" : "false") << "}";
    return s.str();
}

}  // namespace trec_cli
