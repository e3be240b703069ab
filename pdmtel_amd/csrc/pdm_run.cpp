// pdm_run -- command-line driver replacing the reference's main.cpp menu
// (main.cpp:15-85): option 1 runs InitialProblem + SolveDynamic, option 2
// InitialProblem + SolveStatic, option 4 InitialProblem + PostAFD, option 5
// InitialProblem + PostFormatData (both read pds_tecplot_<step>.dat and ask on
// stdin), anything else quits with the reference's message.  The working
// directory (which holds pds_input.txt) is argv[1] instead of the reference's
// stdin prompt (pdDataManager3.cpp:1503-1541); the option is argv[2], or read
// from stdin as the reference's SelectMenu does.
//
//   pdm_run <wkdir> [1|2|4|5] [--fp32c] [--fp32c2] [--device N]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "pdm_facade.hpp"

namespace {

int select_menu()
{
    std::cout << " Please choose an option to run the program: \n"
              << "   - Solve the problem using dynamic solution (1) \n"
              << "   - Solve the problem using static solution (2) \n"
              << "   - Post: Read output file and calculate AFD on given node (4) \n"
              << "   - Post: Convert output file to formatted data file (5) \n"
              << "   - Any other numbers to exit " << std::endl
              << " Your input: ";
    int i = 0;
    if (!(std::cin >> i)) i = 0;
    return i;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <wkdir> [1|2|4|5] [--fp32c|--fp32c2] [--device N]\n", argv[0]);
        return 2;
    }
    pdm_options opt;
    pdm_options_default(&opt);
    int option = -1;
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--fp32c")) opt.precision = PDM_FP32C;
        else if (!std::strcmp(argv[a], "--fp32c2")) opt.precision = PDM_FP32C2;
        else if (!std::strcmp(argv[a], "--device") && a + 1 < argc) opt.device = std::atoi(argv[++a]);
        else option = std::atoi(argv[a]);
    }
    try {
        pdmtel::pdDataManager dat(argv[1], &opt);
        if (option < 0) option = select_menu();
        const std::string out = dat.GetWkDir() + "pds_output.dat";
        auto t0 = std::chrono::steady_clock::now();
        if (option == 1) {
            pdmtel::pdSolve::InitialProblem(dat);
            pdmtel::pdSolve::SolveDynamic(dat);
        } else if (option == 2) {
            pdmtel::pdSolve::InitialProblem(dat);
            pdm_static_stats st = pdmtel::pdSolve::SolveStatic(dat);
            std::ofstream f(out.c_str(), std::ios::app);
            f << "** Static solution: matrix-free PCG, " << st.iterations << " iterations, " << st.refinements
              << " refinement sweeps, relative residual " << st.rel_residual << " **" << std::endl;
            f.close();
            pdmtel::pdSolve::WriteTecplot(dat, 1);
        } else if (option == 4) {
            pdmtel::pdSolve::InitialProblem(dat);
            pdmtel::pdSolve::PostAFD(dat, std::cin, std::cout);
        } else if (option == 5) {
            pdmtel::pdSolve::InitialProblem(dat);
            pdmtel::pdSolve::PostFormatData(dat, std::cin, std::cout);
        } else {
            std::ofstream f(out.c_str(), std::ios::app);
            f << std::endl << "user quits without doing any calculation!" << std::endl;
            return 0;
        }
        long long ms = (long long)std::chrono::duration_cast<std::chrono::milliseconds>(
                           std::chrono::steady_clock::now() - t0).count();
        long long hour = ms / 3600000, minute = (ms / 1000 - hour * 3600) / 60,
                  second = ms / 1000 - hour * 3600 - minute * 60;
        std::ofstream f(out.c_str(), std::ios::app);
        f << "The total CPU tick = " << ms << " (millisecond)" << std::endl;
        f << "The total CPU time is about " << hour << " hours, " << minute << " minutes, and " << second
          << " seconds" << std::endl;
    } catch (const pdmtel::Error& e) {
        std::fprintf(stderr, "pdm_run: %s (code %d)\n", e.what(), e.code());
        return 1;
    }
    return 0;
}
