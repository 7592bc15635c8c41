// egs_dat.h -- the .egsdat state files of egs++ applications.
//
// EGS_Application::outputData / readData / addState
// (egs++/egs_application.cpp:482-589) write and read, in this order:
//   the run control state   ndone (egsStoreI64), cpu time
//                           (egs_run_control.cpp:149-166)
//   the current case        egsStoreI64
//   the RNG state           count, np, ip, the np buffered numbers, then
//                           EGS_Ranmar's ix jx c iseed1 iseed2 high_res and
//                           u[97] (egs_rndm.cpp:83-131, 265-281)
//   the source state        count, then the spectrum's count, sum_E,
//                           sum_E2 (egs_base_source.h:459-492, 740-760)
//   each ausgab object      for EGS_DoseScoring: the last case, then its
//                           region-dose, medium-dose and dose-file scoring
//                           arrays, each present if the object scores it
//                           (egs_dose_scoring.cpp:481-540); an
//                           EGS_ScoringArray is nreg, current_ncase_short,
//                           current_ncase, current_ncase_65536 and per
//                           region "current_ncase sum sum2"
//                           (egs_scoring.h:144-165, 316-365).
// egsStoreI64 writes n as " n/10^9 n%10^9" (egs_functions.cpp:58-80).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace egshost {

struct EgsScoringArray {
    int nreg = 0;
    int current_ncase_short = 0;
    int64_t current_ncase = 0, current_ncase_65536 = 0;
    std::vector<int64_t> last;          // per region: last case that scored
    std::vector<double> sum, sum2;      // per region
};

struct EgsDat {
    int64_t ndone = 0;
    double cpu = 0;
    int64_t current_case = 0;
    // EGS_Ranmar state
    int64_t rng_count = 0;
    int rng_ip = 0;
    std::vector<double> rarray;
    int ix = 0, jx = 0, c = 0, iseed1 = 0, iseed2 = 0, high_res = 0;
    std::vector<double> u;
    // source and spectrum counters
    int64_t src_count = 0, spec_count = 0;
    double sum_E = 0, sum_E2 = 0;
    // one EGS_DoseScoring object
    int64_t last_case = 0;
    std::vector<EgsScoringArray> arrays;
};

// The ranmar state after seeding with (ixx, jxx) (EGS_Ranmar::setState,
// egs_rndm.cpp:303-336), buffer of 128 numbers exhausted.
void ranmar_initial_state(int ixx, int jxx, EgsDat &d);

void write_egsdat(const std::string &path, const EgsDat &d);
// narrays: the number of scoring arrays the dose-scoring object holds.
// Throws std::runtime_error on a malformed or truncated file.
EgsDat read_egsdat(const std::string &path, int narrays);
// EGS_Application::addState: counters and sums added, RNG state replaced.
void add_egsdat(EgsDat &acc, const EgsDat &d);

}  // namespace egshost
