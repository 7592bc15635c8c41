// dosxyz_input.h -- DOSXYZnrc input files for the host application.
//
// Reads the record-oriented input of DOSXYZnrc
// (user_codes/dosxyznrc/dosxyznrc.mortran:150-330, read at :1520-1900) and
// of its source routines for isource 0, 1 and 3
// (user_codes/dosxyznrc/srcxyznrc.mortran:104-146, 606-660, read at
// :1010-1060 and :1760-1830), followed by the EGSnrc transport parameter
// block (":Start MC Transport Parameter:"), which is handed on as text.
//
// Records are read the way the Fortran edit descriptors of DOSXYZnrc read
// them in practice: values separated by commas and/or blanks, an empty
// field between two commas and a missing trailing field read as zero.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace egshost {

struct DosxyzInput {
    std::string title;
    int nmed = 0;                       // 0: CT phantom from an .egsphant file
    std::vector<std::string> media;     // nmed > 0
    std::string phantom_file;           // nmed == 0
    double ecutin = 0, pcutin = 0;
    int zeroairdose = 0;                // nmed == 0: zero the dose in air
    std::vector<double> xb, yb, zb;     // nmed > 0: voxel boundaries
    // nmed > 0, per voxel (x fastest): medium number 1..nmed (0 = vacuum)
    // and density (0 = the medium's default density)
    std::vector<int> med;
    std::vector<double> rho;
    // source (srcxyznrc)
    int iqin = 0, isource = 0;
    std::vector<double> src;            // the values after iqin, isource
    int enflag = 0;
    double ein = 0;                     // enflag 0: kinetic energy (MeV)
    std::string spectrum_file;          // enflag 1
    // record 13
    int64_t ncase = 0;
    int iwatch = 0;
    double timmax = 0;
    int inseed1 = 0, inseed2 = 0;
    double beam_size = 0;
    int ismooth = 0, irestart = 0, idat = 0, ireject = 0;
    double esave_global = 0;
    int nrcycl = 0, iparallel = 0, parnum = 0, n_split = 0, ihowfarless = 0, i_phsp_out = 0,
        i_bindos = 0;
    std::string transport_text;         // everything after record 13
};

// True if the file reads as DOSXYZnrc input (its second record is the
// number of media) rather than an egs++ input with ":start ...:" blocks.
bool is_dosxyz_input(const std::string &path);

// Parse a DOSXYZnrc input file; throws std::runtime_error with the record
// that could not be read.
DosxyzInput read_dosxyz_input(const std::string &path);

// A DOSXYZnrc/BEAMnrc .spectrum file (srcxyznrc.mortran:1811-1826): title,
// NENSRC, ENMIN, IMODE, then NENSRC (upper edge, weight) pairs; IMODE 0 =
// counts per bin, 1 = counts per MeV.  Returns the inverse-CDF grid of
// srcinit (:1874-1899): ngrid points, c1[k] = CDFINV(k+1,1), c2[k] =
// CDFINV(k+1,2).
void dosxyz_spectrum_cdfinv(const std::string &path, int ngrid, std::vector<double> &c1,
                            std::vector<double> &c2, double *emax);

}  // namespace egshost
