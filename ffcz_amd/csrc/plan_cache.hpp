// FFT plan cache of the autotuner: the tile variants chosen for a (device,
// library build, type, shape, FFT variant, bound mode), kept for the process
// and in a text file shared by processes.
//
// The file (one plan per line, later lines replace earlier ones with the same
// key) is read once per process, on the first lookup, and rewritten under an
// exclusive flock whenever a tuning finishes, so concurrent processes merge
// their plans.  Its path is FFCZ_PLAN_CACHE, or $XDG_CACHE_HOME/ffcz/plans.txt
// (~/.cache/ffcz/plans.txt); an empty FFCZ_PLAN_CACHE, "off" or
// ffcz_set_plan_cache(NULL) keeps plans in the process only.  A plan written
// by another build of the library, or on another device model, never matches.
#pragma once

#include <cstddef>
#include <string>

#include "solver.hpp"

namespace ffcz {

struct PlanKey {
    size_t bytes, nx, ny, nz;
    int prec, mode;
};

enum TuneSource : int { TUNE_NONE = 0, TUNE_TIMED = 1, TUNE_PROCESS = 2, TUNE_DISK = 3 };

// The plan for k on the current device: from the process's plans, else from
// the file.  Returns TUNE_PROCESS, TUNE_DISK or TUNE_NONE (not found).
int plan_lookup(const PlanKey& k, TileVars& v);
// Keeps the plan for the process and merges it into the file (best effort: a
// file that cannot be written leaves the process's plan in place).
void plan_store(const PlanKey& k, const TileVars& v);
// The file's path ("" when plans stay in the process).  Setting it drops the
// process's record of what the previous file held.
std::string plan_cache_path();
void set_plan_cache_path(const char* path);

}  // namespace ffcz
