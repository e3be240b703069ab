// pdm_deck.cpp -- pds_input.txt parser (see pdm_deck.h).
#include "pdm_deck.h"

#include <cstdlib>
#include <sstream>

#include "pdm.h"

namespace pdm {
namespace {

std::string trim(const std::string& s)
{
    const char* ws = " \t\r\n";
    size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return std::string();
    size_t e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

struct KeywordFile {
    std::vector<std::string> lines;

    explicit KeywordFile(const std::string& text)
    {
        std::string cur;
        std::istringstream is(text);
        while (std::getline(is, cur)) lines.push_back(trim(cur));
    }

    // FindKeyword (pdDataManager.h:372-421)
    bool find(const std::string& kw) const
    {
        for (const auto& l : lines)
            if (l == kw) return true;
        return false;
    }

    // Tokens of the line after the first line equal to kw (ReadKeyword, pdDataManager.h:270-365).
    bool values(const std::string& kw, std::vector<std::string>& tok) const
    {
        for (size_t i = 0; i < lines.size(); ++i) {
            if (lines[i] != kw) continue;
            tok.clear();
            if (i + 1 < lines.size()) {
                std::istringstream iss(lines[i + 1]);
                std::string t;
                while (iss >> t) tok.push_back(t);
            }
            return true;
        }
        return false;
    }
};

class Reader {
public:
    Reader(const KeywordFile& f, std::string& err) : f_(f), err_(err) {}

    bool has(const std::string& kw) const { return f_.find(kw); }

    // REQ keyword presence (FindKeyword with REQ, pdDataManager.h:400-403)
    bool require(const std::string& kw)
    {
        if (f_.find(kw)) return true;
        fail("FindKeyword: required keyword " + kw + " not found");
        return false;
    }

    bool read_double(const std::string& kw, int pos, double& v)
    {
        std::string t;
        if (!token(kw, pos, t)) return false;
        char* end = nullptr;
        v = std::strtod(t.c_str(), &end);
        if (end == t.c_str()) return fail("ReadKeyword: bad number '" + t + "', keyword = " + kw);
        return true;
    }

    // an int read by operator>> stops at the first non-digit ("2.5" reads 2)
    bool read_int(const std::string& kw, int pos, int& v)
    {
        std::string t;
        if (!token(kw, pos, t)) return false;
        char* end = nullptr;
        long x = std::strtol(t.c_str(), &end, 10);
        if (end == t.c_str()) return fail("ReadKeyword: bad integer '" + t + "', keyword = " + kw);
        v = (int)x;
        return true;
    }

    bool read_bool(const std::string& kw, int pos, bool& v)
    {
        int x = 0;
        if (!read_int(kw, pos, x)) return false;
        v = (x != 0);
        return true;
    }

    bool read_string(const std::string& kw, int pos, std::string& v) { return token(kw, pos, v); }

    bool fail(const std::string& m)
    {
        if (err_.empty()) err_ = m;
        return false;
    }

private:
    bool token(const std::string& kw, int pos, std::string& out)
    {
        std::vector<std::string> tok;
        if (!f_.values(kw, tok)) return fail("ReadKeyword: keyword " + kw + " not found");
        if (pos <= 0 || (size_t)pos > tok.size())
            return fail("ReadKeyword: value missing at desired position, keyword = " + kw);
        out = tok[pos - 1];
        return true;
    }

    const KeywordFile& f_;
    std::string& err_;
};

std::string numbered(const std::string& base, int i) { return base + std::to_string(i); }

bool read_space(Reader& r, const std::string& kw, bool with_ref, Space& s)
{
    if (!r.read_int(kw, 1, s.geo)) return false;
    int p = 2;
    if (with_ref) {
        int ref = 0;
        if (!r.read_int(kw, p++, ref)) return false;
        s.ref = ref - 1;  // ids in code start from zero
    }
    int cnt = (s.geo == 1) ? 6 : (s.geo == 2 ? 5 : 0);
    if (cnt == 0) return r.fail("unsupported space geometry in " + kw);
    for (int k = 0; k < cnt; ++k)
        if (!r.read_double(kw, p++, s.b[k])) return false;
    return true;
}

}  // namespace

int parse_deck(const std::string& text, Deck& d, std::string& err)
{
    KeywordFile f(text);
    Reader r(f, err);
    err.clear();
#define REQ(kw)                              \
    do {                                     \
        if (!r.require(kw)) return PDM_E_PARSE; \
    } while (0)
#define CHK(x)                          \
    do {                                \
        if (!(x)) return PDM_E_PARSE;   \
    } while (0)

    // 1. basic setup (pdDataManager.cpp:11-111)
    if (r.has("problem_title")) CHK(r.read_string("problem_title", 1, d.title));
    if (r.has("grid_center")) {
        CHK(r.read_double("grid_center", 1, d.cx));
        CHK(r.read_double("grid_center", 2, d.cy));
        CHK(r.read_double("grid_center", 3, d.cz));
    }
    REQ("grid_dimension");
    CHK(r.read_int("grid_dimension", 1, d.nx));
    CHK(r.read_int("grid_dimension", 2, d.ny));
    CHK(r.read_int("grid_dimension", 3, d.nz));
    REQ("grid_spacing");
    CHK(r.read_double("grid_spacing", 1, d.dx));
    if (r.has("horizon")) CHK(r.read_double("horizon", 1, d.horizon));
    else d.horizon = 3.015 * d.dx;
    REQ("total_time");
    CHK(r.read_double("total_time", 1, d.total_time));
    REQ("total_step");
    CHK(r.read_int("total_step", 1, d.total_step));
    REQ("step_dump_frequency");
    CHK(r.read_int("step_dump_frequency", 1, d.dump));
    if (r.has("safety_factor")) CHK(r.read_double("safety_factor", 1, d.safety));
    else d.safety = 0.8;
    d.fixed_dt_on = false;
    if (r.has("fixed_time_step")) {
        d.fixed_dt_on = true;
        CHK(r.read_double("fixed_time_step", 1, d.fixed_dt));
    }

    // 2. materials (pdDataManager.cpp:113-179)
    REQ("number_of_materials");
    int nm = 0;
    CHK(r.read_int("number_of_materials", 1, nm));
    for (int i = 0; i < nm; ++i) {
        std::string kw = numbered("material_", i + 1);
        Material m;
        CHK(r.read_int(kw, 1, m.type));
        if (m.type == 1) {
            CHK(r.read_double(kw, 2, m.E));
            CHK(r.read_double(kw, 3, m.sigy));
            CHK(r.read_double(kw, 4, m.ecrit));
            CHK(r.read_double(kw, 5, m.rho));
            m.nu = 0.25;  // pdMaterialBond.cpp:10
        } else if (m.type == 2) {
            CHK(r.read_double(kw, 2, m.E));
            CHK(r.read_double(kw, 3, m.sigy));
            CHK(r.read_double(kw, 4, m.nu));
            CHK(r.read_double(kw, 5, m.ecrit));
            CHK(r.read_double(kw, 6, m.rho));
        } else {
            r.fail("unsupported material type in " + kw);
            return PDM_E_PARSE;
        }
        std::string gk = numbered("gravity_", i + 1);
        if (r.has(gk)) {
            CHK(r.read_double(gk, 1, m.g[0]));
            CHK(r.read_double(gk, 2, m.g[1]));
            CHK(r.read_double(gk, 3, m.g[2]));
        }
        d.mats.push_back(m);
    }

    // 3. boundary conditions (pdDataManager.cpp:181-291)
    REQ("number_of_boundary_conditions");
    int nb = 0;
    CHK(r.read_int("number_of_boundary_conditions", 1, nb));
    for (int i = 0; i < nb; ++i) {
        std::string kw = numbered("boundary_condition_", i + 1);
        BoundaryCondition b;
        CHK(r.read_int(kw, 1, b.type));
        if (b.type < 1 || b.type > 4) {
            r.fail("unsupported boundary condition type in " + kw);
            return PDM_E_PARSE;
        }
        CHK(r.read_int(kw, 2, b.f[0]));
        CHK(r.read_int(kw, 3, b.f[1]));
        CHK(r.read_int(kw, 4, b.f[2]));
        int nv = (b.type == 3) ? 9 : 3;
        for (int k = 0; k < nv; ++k) CHK(r.read_double(kw, 5 + k, b.v[k]));
        CHK(r.read_double(kw, 5 + nv, b.t_end));
        CHK(r.read_double(kw, 6 + nv, b.load));
        d.bcs.push_back(b);
    }

    // 4-6. spaces (pdDataManager.cpp:293-544)
    REQ("number_of_material_spaces");
    int nms = 0;
    CHK(r.read_int("number_of_material_spaces", 1, nms));
    for (int i = 0; i < nms; ++i) {
        Space s;
        CHK(read_space(r, numbered("material_space_", i + 1), true, s));
        if (s.ref < 0 || s.ref >= (int)d.mats.size()) {
            r.fail("material_space_" + std::to_string(i + 1) + " names a missing material");
            return PDM_E_PARSE;
        }
        d.mat_spaces.push_back(s);
    }
    if (r.has("number_of_deletion_spaces")) {
        int nds = 0;
        CHK(r.read_int("number_of_deletion_spaces", 1, nds));
        for (int i = 0; i < nds; ++i) {
            Space s;
            CHK(read_space(r, numbered("deletion_space_", i + 1), false, s));
            d.del_spaces.push_back(s);
        }
    }
    REQ("number_of_boundary_spaces");
    int nbs = 0;
    CHK(r.read_int("number_of_boundary_spaces", 1, nbs));
    for (int i = 0; i < nbs; ++i) {
        Space s;
        CHK(read_space(r, numbered("boundary_space_", i + 1), true, s));
        if (s.ref < 0 || s.ref >= (int)d.bcs.size()) {
            r.fail("boundary_space_" + std::to_string(i + 1) + " names a missing boundary condition");
            return PDM_E_PARSE;
        }
        d.bc_spaces.push_back(s);
    }

    // 7. parameters and integration-mode cascade (pdDataManager.cpp:546-730)
    REQ("damping_coeff");
    CHK(r.read_double("damping_coeff", 1, d.damping));
    REQ("bdryEff_on");
    CHK(r.read_bool("bdryEff_on", 1, d.bdryeff_on));
    d.tol = 1.0e-4 * d.dx;
    REQ("cci_on");
    CHK(r.read_bool("cci_on", 1, d.cci_on));
    if (d.cci_on) {
        d.fgi_on = d.mls_on = d.ai_on = d.err_on = false;
        d.eps = 0.0;
        d.intg_point = 0;
        CHK(r.read_bool("cci_on", 2, d.cci_mdf));
    } else {
        REQ("fgi_on");
        CHK(r.read_bool("fgi_on", 1, d.fgi_on));
        if (d.fgi_on) {
            if (r.has("intg_point")) {
                CHK(r.read_int("intg_point", 1, d.intg_point));
                d.mls_on = (d.intg_point != 1);
            } else {
                d.intg_point = 3;
                d.mls_on = true;
            }
            if (r.has("support_domain_size")) CHK(r.read_double("support_domain_size", 1, d.rw));
            else d.rw = 3.0 * d.dx;
            d.ai_on = d.err_on = false;
            d.eps = 0.0;
            d.cci_mdf = false;
        } else {
            d.cci_mdf = false;
            REQ("ai_on");
            CHK(r.read_bool("ai_on", 1, d.ai_on));
            if (!d.ai_on) {
                r.fail("ReadInputFile: no integration method is used!");
                return PDM_E_PARSE;
            }
            CHK(r.read_bool("ai_on", 2, d.mls_on));
            if (r.has("support_domain_size")) CHK(r.read_double("support_domain_size", 1, d.rw));
            else d.rw = 3.0 * d.dx;
            if (r.has("intg_point")) CHK(r.read_int("intg_point", 1, d.intg_point));
            else d.intg_point = 3;
            if (r.has("error_control_on")) {
                CHK(r.read_bool("error_control_on", 1, d.err_on));
                if (d.err_on) {
                    if (r.has("error_control_eps")) CHK(r.read_double("error_control_eps", 1, d.eps));
                    else d.eps = 0.01;
                }
            }
        }
    }
    if (r.has("grid_file")) CHK(r.read_string("grid_file", 1, d.grid_file));

    // extensions
    if (r.has("grid_jitter")) {
        CHK(r.read_double("grid_jitter", 1, d.jitter));
        int seed = 0;
        CHK(r.read_int("grid_jitter", 2, seed));
        d.seed = (uint64_t)(int64_t)seed;
    }
    if (r.has("number_of_slits")) {
        int ns = 0;
        CHK(r.read_int("number_of_slits", 1, ns));
        for (int i = 0; i < ns; ++i) {
            std::string kw = numbered("slit_", i + 1);
            Slit s;
            CHK(r.read_int(kw, 1, s.type));
            int cnt = (s.type == 1) ? 6 : (s.type == 2 ? 11 : 0);
            if (cnt == 0) {
                r.fail("unsupported slit type in " + kw);
                return PDM_E_PARSE;
            }
            for (int k = 0; k < cnt; ++k) CHK(r.read_double(kw, 2 + k, s.p[k]));
            d.slits.push_back(s);
        }
    }
    if (r.has("traction_as_body_force")) CHK(r.read_bool("traction_as_body_force", 1, d.traction_body));
#undef REQ
#undef CHK
    if (d.nx < 0 || d.ny < 0 || d.nz < 0 || !(d.dx > 0) || !(d.horizon > 0)) {
        err = "grid_dimension / grid_spacing / horizon out of range";
        return PDM_E_PARSE;
    }
    return PDM_OK;
}

}  // namespace pdm
