// pdm_facade.hpp -- C++ facade over the C ABI of pdm.h with the reference
// driver's class and method names, so a caller of pdSolve / pdDataManager
// (main.cpp:21-59) switches by changing its includes and namespace.
//
//   pdmtel::pdDataManager dat(wkdir);           // pdDataManager + SetWkDir + ReadInputFile
//   pdmtel::pdSolve::InitialProblem(dat);       // pdSolve.cpp:10-22
//   pdmtel::pdSolve::SolveDynamic(dat);         // pdSolve.cpp:24-130
//   pdmtel::pdSolve::SolveStatic(dat);          // pdSolve.cpp:132-193
//
// The reference reports errors by exit(0) (pdDataManager.h:400-419); the facade
// throws pdmtel::Error carrying the PDM_E* code and message instead.
#ifndef PDM_FACADE_HPP
#define PDM_FACADE_HPP

#include <cstdint>
#include <stdexcept>
#include <fstream>
#include <istream>
#include <ostream>
#include <string>
#include <vector>

#include "pdm.h"

namespace pdmtel {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

class pdDataManager {
public:
    // Working directory holding pds_input.txt (SetWkDir, pdDataManager3.cpp:1503-1541),
    // parsed at once (ReadInputFile, pdDataManager.cpp:4-758).
    explicit pdDataManager(const std::string& wkdir, const pdm_options* opt = nullptr) : wkdir_(wkdir)
    {
        if (!wkdir_.empty() && wkdir_.back() != '/') wkdir_ += '/';
        pdm_options o;
        if (opt) o = *opt;
        else pdm_options_default(&o);
        check(pdm_create((wkdir_ + "pds_input.txt").c_str(), &o, &h_), nullptr);
        check(pdm_get_problem_info(h_, &info_));
    }
    ~pdDataManager() { pdm_destroy(h_); }
    pdDataManager(const pdDataManager&) = delete;
    pdDataManager& operator=(const pdDataManager&) = delete;

    // InitNode / InitBond / InitMaterial / InitBdryEffectFactor in one call.
    void Init() { check(pdm_init(h_)); inited_ = true; }
    bool IsInitialised() const { return inited_; }

    // Getters of pdDataManager.h:104-142.
    std::string GetWkDir() const { return wkdir_; }
    int GetRank() const { return 0; }
    int GetNumNode() const { return (int)pdm_num_nodes(h_); }
    int64_t GetNumBond() const { return pdm_num_bonds(h_); }
    int GetNumMaterial() const { return info_.num_materials; }
    int GetNumBC() const { return info_.num_bcs; }
    int GetNumMatSpace() const { return info_.num_material_spaces; }
    int GetNumDelSpace() const { return info_.num_deletion_spaces; }
    int GetNumBdSpace() const { return info_.num_boundary_spaces; }
    int GetStepDumpFrequency() const { return info_.step_dump_frequency; }
    int GetGridDimensionX1() const { return info_.grid_dimension[0]; }
    int GetGridDimensionX2() const { return info_.grid_dimension[1]; }
    int GetGridDimensionX3() const { return info_.grid_dimension[2]; }
    int GetTotalStep() const { return info_.total_step; }
    double GetHorizon() const { return info_.horizon; }
    double GetGridCenterX1() const { return info_.grid_center[0]; }
    double GetGridCenterX2() const { return info_.grid_center[1]; }
    double GetGridCenterX3() const { return info_.grid_center[2]; }
    double GetGridSpacing() const { return info_.grid_spacing; }
    double GetTotalTime() const { return info_.total_time; }
    bool IsFixedTimeStepOn() const { return info_.fixed_time_step > 0.0; }
    bool IsMLSOn() const { return info_.mls_on != 0; }
    // CalcStableTimeStep (pdDataManager2.cpp:6-10)
    double CalcStableTimeStep() const
    {
        double dt = 0.0;
        check(pdm_get_timestep(h_, &dt, nullptr));
        return dt;
    }
    // GetKinetTotal / GetElastStep (pdDataManager3.cpp:1415-1428), of the last step
    double GetKinetTotal() const
    {
        double ke = 0.0;
        check(pdm_get_energy(h_, &ke, nullptr));
        return ke;
    }
    double GetElastStep() const
    {
        double ee = 0.0;
        check(pdm_get_energy(h_, nullptr, &ee));
        return ee;
    }
    // Node data in reference order (pdNode getters, pdNode.h:27-142): 3N each.
    void GetDisplacement(std::vector<double>& U, std::vector<double>& V) const
    {
        U.resize(3 * (size_t)GetNumNode());
        V.resize(U.size());
        check(pdm_get_state(h_, U.data(), V.data()));
    }
    void GetPosition(std::vector<double>& X) const
    {
        X.resize(3 * (size_t)GetNumNode());
        check(pdm_get_nodes(h_, X.data(), nullptr, nullptr));
    }

    // The post-processing readers (PostAFD / PostFormatData, pdDataManager3.cpp:940-1120): a
    // tecplot dump's first four lines are skipped, then per node id, X, U, V, material, BC,
    // bdryEff and damage; U and V become the state, bdryEff replaces the factors of the AFD,
    // the damage column is read and dropped as the reference does.
    void ReadTecplot(std::istream& in)
    {
        std::string line;
        for (int k = 0; k < 4; ++k) std::getline(in, line);
        const int n = GetNumNode();
        std::vector<double> U(3 * (size_t)n), V(3 * (size_t)n);
        dump_be_.assign((size_t)n, 1.0);
        for (int k = 0; k < n; ++k) {
            long long id;
            int mid, bid;
            double x[3], u[3], v[3], be, dmg;
            if (!(in >> id >> x[0] >> x[1] >> x[2] >> u[0] >> u[1] >> u[2] >> v[0] >> v[1] >> v[2] >> mid >> bid >>
                  be >> dmg) || id < 0 || id >= n)
                throw Error(PDM_E_PARSE, "tecplot dump ends early or holds a bad node id");
            for (int c = 0; c < 3; ++c) {
                U[3 * id + c] = u[c];
                V[3 * id + c] = v[c];
            }
            dump_be_[id] = be;
            (void)dmg;
        }
        check(pdm_set_state(h_, U.data(), V.data()));
    }
    // CalcArealForceDensity (pdDataManager3.cpp:442-520) at the given nodes: 9 values each.
    void CalcArealForceDensity(const std::vector<int64_t>& nodes, std::vector<double>& sigma) const
    {
        sigma.resize(9 * nodes.size());
        check(pdm_areal_force_density(h_, nodes.data(), (int64_t)nodes.size(),
                                      dump_be_.empty() ? nullptr : dump_be_.data(), sigma.data()));
    }
    const std::vector<double>& DumpBdryEff() const { return dump_be_; }

    pdm_handle* handle() const { return h_; }
    const pdm_problem_info& info() const { return info_; }

    void check(int rc) const { check(rc, h_); }

private:
    static void check(int rc, const pdm_handle* h)
    {
        if (rc != PDM_OK) throw Error(rc, pdm_last_error(h));
    }

    std::string wkdir_;
    pdm_handle* h_ = nullptr;
    pdm_problem_info info_{};
    bool inited_ = false;
    std::vector<double> dump_be_;  // bdryEff read by ReadTecplot
};

struct pdSolve {
    // InitialProblem (pdSolve.cpp:10-22)
    static void InitialProblem(pdDataManager& dat) { dat.Init(); }
    // SolveDynamic (pdSolve.cpp:24-130): pds_output.dat, pds_tecplot_<step>.dat in the working dir
    static void SolveDynamic(pdDataManager& dat) { dat.check(pdm_solve_dynamic(dat.handle(), dat.GetWkDir().c_str())); }
    // SolveStatic (pdSolve.cpp:132-193), as matrix-free PCG to rel_tol; the solution is the state U
    static pdm_static_stats SolveStatic(pdDataManager& dat, double rel_tol = 1e-10, int max_iterations = 0)
    {
        pdm_static_stats st{};
        dat.check(pdm_solve_static(dat.handle(), rel_tol, max_iterations, &st));
        return st;
    }
    // WriteTecplot (pdSolve.cpp:708-732)
    static void WriteTecplot(const pdDataManager& dat, int step)
    {
        std::string path = dat.GetWkDir() + "pds_tecplot_" + std::to_string(step) + ".dat";
        dat.check(pdm_write_tecplot(dat.handle(), path.c_str(), step));
    }

    // PostAFD (pdSolve.cpp:810-834, pdDataManager3.cpp:1101-1220): read pds_tecplot_<step>.dat,
    // then the areal force density on a coordinate plane or at one node, into pds_AFD_<step>.dat,
    // with the reference's prompts on `ask` / `in`.
    static void PostAFD(pdDataManager& dat, std::istream& in, std::ostream& ask)
    {
        int step = 0;
        ask << " Please give the step number you want to read: ";
        in >> step;
        std::ifstream dump((dat.GetWkDir() + "pds_tecplot_" + std::to_string(step) + ".dat").c_str());
        if (!dump) throw Error(PDM_E_IO, "ReadTecplotFile: pdMTEL tecplot output file not exist!");
        dat.ReadTecplot(dump);
        std::ofstream out((dat.GetWkDir() + "pds_AFD_" + std::to_string(step) + ".dat").c_str());
        std::vector<double> X;
        dat.GetPosition(X);
        const int n = dat.GetNumNode();
        for (;;) {
            ask << " Choose an option to calculate the areal force density: \n"
                << " - AFD at every node on a cross-section plane (1) \n"
                << " - AFD at single node (2) \n"
                << " Your input: ";
            int opt = 0;
            in >> opt;
            std::vector<double> sig;
            if (opt == 1) {
                int norm = 0;
                double coord = 0.0;
                ask << " Give the normal direction of the plane (1, 2, or 3): ";
                in >> norm;
                ask << " Give the location of the plane: ";
                in >> coord;
                out << " norm = " << norm << ", location = " << coord << std::endl;
                std::vector<int64_t> ids;
                for (int i = 0; i < n; ++i)
                    if (norm >= 1 && norm <= 3 && X[3 * i + norm - 1] == coord) ids.push_back(i);
                dat.CalcArealForceDensity(ids, sig);
                for (size_t k = 0; k < ids.size(); ++k) {
                    const int64_t i = ids[k];
                    out << i << "\t" << X[3 * i] << "\t" << X[3 * i + 1] << "\t" << X[3 * i + 2];
                    for (int q = 0; q < 9; ++q) out << "\t" << sig[9 * k + q];
                    out << "\n";
                }
            } else if (opt == 2) {
                long long id = 0;
                ask << " Give the node id: ";
                in >> id;
                if (id < 0 || id >= n) throw Error(PDM_E_ARG, "node id out of range");
                dat.CalcArealForceDensity({(int64_t)id}, sig);
                out << " AFD at node " << X[3 * id] << "\t" << X[3 * id + 1] << "\t" << X[3 * id + 2] << std::endl;
                out << " sigma11 = " << sig[0] << " sigma12 = " << sig[1] << " sigma13 = " << sig[2] << std::endl
                    << " sigma21 = " << sig[3] << " sigma22 = " << sig[4] << " sigma23 = " << sig[5] << std::endl
                    << " sigma31 = " << sig[6] << " sigma32 = " << sig[7] << " sigma33 = " << sig[8] << std::endl
                    << std::endl;
            }
            ask << " Another run? (y/n) ";
            std::string run;
            if (!(in >> run) || run == "n") break;
        }
    }

    // PostFormatData (pdSolve.cpp:782-808, pdDataManager3.cpp:940-1099): read pds_tecplot_<step>.dat
    // and write node lines of a plane or one node, or the normal strain at a plane's nodes from
    // the family member one spacing further along the normal, into pds_format_<step>.dat.
    static void PostFormatData(pdDataManager& dat, std::istream& in, std::ostream& ask)
    {
        int step = 0;
        ask << " Please give the step number you want to read: ";
        in >> step;
        std::ifstream dump((dat.GetWkDir() + "pds_tecplot_" + std::to_string(step) + ".dat").c_str());
        if (!dump) throw Error(PDM_E_IO, "ReadTecplotFile: pdMTEL tecplot output file not exist!");
        dat.ReadTecplot(dump);
        std::ofstream out((dat.GetWkDir() + "pds_format_" + std::to_string(step) + ".dat").c_str());
        const int n = dat.GetNumNode();
        std::vector<double> X, U, V;
        dat.GetPosition(X);
        dat.GetDisplacement(U, V);
        std::vector<int32_t> mat(n), bc(n);
        dat.check(pdm_get_nodes(dat.handle(), nullptr, mat.data(), bc.data()));
        const std::vector<double>& be = dat.DumpBdryEff();
        // pdNode::Print (pdNode.cpp:418-433); the model just read has broken no bond, so its
        // damage fraction (pdNode.cpp:225-232) is 0
        auto node_line = [&](int i) {
            out << i << "\t" << X[3 * i] << "\t" << X[3 * i + 1] << "\t" << X[3 * i + 2] << "\t" << U[3 * i] << "\t"
                << U[3 * i + 1] << "\t" << U[3 * i + 2] << "\t" << V[3 * i] << "\t" << V[3 * i + 1] << "\t"
                << V[3 * i + 2] << "\t" << mat[i] << "\t" << bc[i] << "\t" << be[i] << "\t" << 0.0;
        };
        const int64_t E = pdm_num_entries(dat.handle());
        std::vector<int64_t> rp(n + 1);
        std::vector<int32_t> col(E > 0 ? E : 1);
        dat.check(pdm_get_csr(dat.handle(), rp.data(), col.data(), nullptr));
        const double dx = dat.GetGridSpacing();
        for (;;) {
            ask << " Choose a format option: \n"
                << " - Nodal data at every node on a cross-section plane (1) \n"
                << " - Nodal data at single node (2) \n"
                << " - Nodal strain at every node on a cross-section plane (3) \n"
                << " Your input: ";
            int opt = 0;
            in >> opt;
            if (opt == 1 || opt == 3) {
                int norm = 0;
                double coord = 0.0;
                ask << " Give the normal direction of the plane (1, 2, or 3): ";
                in >> norm;
                ask << " Give the location of the plane: ";
                in >> coord;
                out << " norm = " << norm << ", location = " << coord << std::endl;
                for (int i = 0; i < n; ++i) {
                    if (norm < 1 || norm > 3 || !(X[3 * i + norm - 1] == coord)) continue;
                    if (opt == 1) {
                        node_line(i);
                        out << std::endl;
                        continue;
                    }
                    out << i << "\t" << X[3 * i] << "\t" << X[3 * i + 1] << "\t" << X[3 * i + 2];
                    const int c = norm - 1;
                    for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {  // the member one spacing along e_c
                        const int j = col[e];
                        bool at = X[3 * j + c] == X[3 * i + c] + dx;
                        for (int q = 0; q < 3; ++q)
                            if (q != c) at = at && X[3 * j + q] == X[3 * i + q];
                        if (at) {
                            out << "\t" << (U[3 * j + c] - U[3 * i + c]) / (X[3 * j + c] - X[3 * i + c]);
                            break;
                        }
                    }
                    out << std::endl;
                }
                out << std::endl;
            } else if (opt == 2) {
                long long id = 0;
                ask << " Give the node id: ";
                in >> id;
                if (id < 0 || id >= n) throw Error(PDM_E_ARG, "node id out of range");
                node_line((int)id);
                out << std::endl;
            }
            ask << " Another run? (y/n) ";
            std::string run;
            if (!(in >> run) || run == "n") break;
        }
    }
};

}  // namespace pdmtel

#endif
