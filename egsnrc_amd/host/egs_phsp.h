// egs_phsp.h -- EGSnrc phase-space files (MODE0 / MODE2) for the host
// application: the record layout read by EGS_PhspSource
// (egs++/sources/egs_phsp_source/egs_phsp_source.cpp:79-431).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

struct EgsPhsp {
    bool mode2 = false;
    int64_t n = 0;             // particles in the header
    int64_t nphot = 0;         // photons in the header
    float emax = 0, emin = 0;  // kinetic energy range in the header
    float pinc = 0;            // incident particles of the simulation that wrote it
    std::vector<int32_t> latch;
    std::vector<float> e;      // kinetic energy (MeV)
    std::vector<float> x, y, u, v, wt;   // wt carries the sign of w
    std::vector<int8_t> q;
    std::vector<uint8_t> new_history;    // the record started a new primary history
};

// Read a whole MODE0/MODE2 file (either byte order); throws
// std::runtime_error.  Charged particles' total energies are converted to
// kinetic (readParticle, :423-425).
EgsPhsp read_egsphsp(const std::string &path);

// Read an IAEA phase space (base.IAEAheader + base.IAEAphsp), the format of
// the IAEA library the reference's iaea_phsp_source reads
// (iaea_phsp/iaea_header.cpp:90-180 header blocks, iaea_record.cpp:157-230
// records; charge mapping iaea_phsp_source.cpp:372-384): per record a
// signed char type (1 photon, 2 electron, 3 positron; its sign is the sign
// of w), then floats E (negative: a new history), and x, y, z, u, v, weight
// where the header's RECORD_CONTENTS stores them (RECORD_CONSTANT gives the
// others), extra floats, extra int32s.  Energies are kinetic.  The
// particles must lie on one z plane, returned in zplane.
EgsPhsp read_iaea_phsp(const std::string &base, double &zplane);
