// opts.h -- TiltRec command line (src/TiltRec/src/opts/opts.h:17-38,
// opts.cpp:126-360) for the HIP CLI.
#pragma once

#include <string>

namespace trec_cli {

struct Options {
    std::string input, output, initial, angle;
    // --geometry offset,pitch,zshift,thickness
    float offset = 0, pitch = 0, zshift = 0;
    int thickness = 0;
    // --method NAME[,iterations,gamma] or ADMM,iterations,cg,gamma,soft
    std::string method;
    int iteration = 0, cgiter = 0;
    float gamma = 1.0f, soft = 0;
    std::string axis = "y";
    bool f2b = false;
    // additions beyond the reference's options
    std::string rules;     // "mpi" | "cuda" ("" = follow the program name)
    std::string storage = "f32";
    int device = 0;
    int output_mode = 2;   // MRC mode of the volume: 2 (float) or 12 (half)
    bool mrc2014 = false;  // MRC2014 header: MAP stamp, machine stamp, sampling, whole-map statistics
    bool dry_run = false;  // print the parsed options as JSON and exit
};

void usage();
// 1: run, 0: help printed, -1: invalid command line (message on stderr).
int get_opts(int argc, char **argv, Options &o);
std::string opts_json(const Options &o);

}  // namespace trec_cli
