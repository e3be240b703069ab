// pdm_deck.h -- pds_input.txt deck model and parser.
//
// Mirrors pdDataManager::ReadInputFile (pdDataManager.cpp:4-758) and its
// keyword reader (ReadKeyword / FindKeyword, pdDataManager.h:270-421): a
// keyword is a line equal to the keyword after trimming, its values are the
// whitespace-separated tokens of the next line, the first matching line wins.
// Differences: CR and tabs are trimmed too (the reference trims only ' ', so
// a CRLF deck matches no keyword on Linux), and errors are returned instead of
// exit(0).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pdm {

struct Material {
    int type = 1;                 // 1 bond-based (PMB), 2 state-based
    double E = 0, sigy = 0, nu = 0.25, ecrit = 0, rho = 0;
    double g[3] = {0, 0, 0};      // gravity_N
};

struct BoundaryCondition {
    int type = 0;                 // 1 DISPL, 2 VELOC, 3 DISPLGRAD, 4 TRACTION
    int f[3] = {0, 0, 0};
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double t_end = 0, load = 0;
};

struct Space {
    int geo = 1;                  // 1 cuboid (b = x1lo x1hi x2lo x2hi x3lo x3hi), 2 cylinder (b = r cx cy x3lo x3hi)
    int ref = -1;                 // material / bc index, 0-based
    double b[6] = {0, 0, 0, 0, 0, 0};
};

struct Slit {
    int type = 1;                 // 1 axis-normal rectangle (box with one zero extent), 2 oriented rectangle
    double p[11] = {0};           // 2: centre(3) normal(3) in-plane axis(3) half_t half_s
};

struct Deck {
    std::string title = "Peridynamic test problem";
    double cx = 0, cy = 0, cz = 0;
    int nx = 0, ny = 0, nz = 0;
    double dx = 0, horizon = 0;
    double total_time = 0;
    int total_step = 0, dump = 0;
    double safety = 0.8;
    bool fixed_dt_on = false;
    double fixed_dt = 0;
    std::vector<Material> mats;
    std::vector<BoundaryCondition> bcs;
    std::vector<Space> mat_spaces, del_spaces, bc_spaces;
    double damping = 0;
    bool bdryeff_on = true;
    bool cci_on = false, cci_mdf = false, fgi_on = false, mls_on = false, ai_on = false, err_on = false;
    int intg_point = 0;
    double rw = 0, eps = 0;
    double tol = 0;               // 1e-4 dx (pdDataManager.cpp:569)
    std::string grid_file;        // optional external grid (pdDataManager.cpp:831-841)
    // extensions (no reference counterpart)
    double jitter = 0;            // grid_jitter: amplitude as a fraction of dx
    uint64_t seed = 0;
    std::vector<Slit> slits;      // number_of_slits / slit_N
    bool traction_body = false;   // traction_as_body_force
};

// Parse deck text.  Returns 0 or a PDM_E* code with a message in err.
int parse_deck(const std::string& text, Deck& deck, std::string& err);

}  // namespace pdm
