// bin_fednl_local_pp: drop-in executable for dopt's single-node FedNL-PP
// simulation (dopt/bin_fednl_local_pp/src/main.cpp:22-56); all work runs on
// the GPU engine.
#include <chrono>
#include <iostream>

#include "fednl_hip.h"

int main(int argc, char** argv)
{
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = fednl_cli_pp_main(argc, argv);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::cout << '\n' << "Time spent to execution is " << ms << " milliseconds\n" << '\n';
    return rc;
}
