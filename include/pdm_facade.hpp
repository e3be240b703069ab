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
};

}  // namespace pdmtel

#endif
