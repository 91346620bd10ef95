// cosenu -- command-line driver with the reference's behaviour
// (src/main.cpp:58-231), running the time step on the GPU.
//
//   cosenu --conf job.config [--ff] [--mode strict|fast] [--device N]
//
// Same inputs (job.config, and with --ff the files G0.bin, state.bin or
// state0.bin), same outputs (state0.bin, G0.bin, rho_<t>.dat,
// conserved_quantities.dat, survival_probability.dat, time.txt) and the same
// schedule: N_ITER - 1 RK4 steps from t0 = 1 (or from the iteration stored in
// state.bin), analysis every ANAL_EVERY, dumps every dump_interval.
// Differences, on purpose: failures exit with status 1 and a message instead
// of exit(0) (parser.hpp:107-113, 220-243); N_ITER < 10 does not divide by
// zero (main.cpp:195); with `checkpoint_every : N` in job.config the state is
// written to state.bin every N steps (the write the reference leaves commented
// out at main.cpp:201), stamped with the next iteration so that `--ff` resumes
// the run exactly.  Runs of steps between outputs are handed to the library in
// one call so they replay as captured HIP graphs.
#include <chrono>
#include <cstring>
#include <iostream>
#include <string>

#include "nuosc.hpp"
#include "params.hpp"

using namespace cosenu_host;

static void usage()
{
    std::cout << "[ FAIL ]...The configuration should be passed on runtime." << std::endl
              << "Use --conf to specify them." << std::endl
              << "General format:\n\t$./cosenu --conf <Configuration file name>\n\n"
              << "If loading field variables from .bin file:\n"
              << "\t$./cosenu --conf <Configuration file name> --ff\n"
              << "The --ff flag will look for binary files named state.bin and G0.bin in the "
                 "same folder.\n";
}

int main(int argc, char *argv[])
{
    std::string conf;
    bool ff = false;
    int device = 0;
    int mode_override = -1;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--conf") && i + 1 < argc)
            conf = argv[++i];
        else if (!strcmp(argv[i], "--ff"))
            ff = true;
        else if (!strcmp(argv[i], "--device") && i + 1 < argc)
            device = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--mode") && i + 1 < argc) {
            std::string m = argv[++i];
            mode_override = (m == "fast") ? COSENU_FAST : COSENU_STRICT;
        }
    }
    if (conf.empty()) {
        usage();
        return 1;
    }
    if (ff)
        std::cout << "Initializing from file.\n";

    Params pars;
    std::string err = pars.load(conf);
    if (!err.empty()) {
        std::cout << err << std::endl << "Exiting." << std::endl;
        return 1;
    }
    if (mode_override >= 0)
        pars.mode = mode_override;

    try {
        NuOsc state(pars, device);
        const int N_ITER = pars.N_ITER;
        std::cout << std::setw(30) << "NUMBER OFITERATIONS: " << N_ITER << std::endl;

        int t0 = 0;
        if (!ff) {
            state.initialize(); // also copy_state(v_stat -> v_stat0)
            state.write_state0();
            t0 = 1;
        } else {
            state.read_G0();
            t0 = state.read_state();
            state.read_state0();
        }
        std::cout << "Starting time = " << t0 << "\n";

        if (!ff) {
            state.analyse(0);
            state.surv_prob(0);
            state.dump_rho(0);
        }

        auto start = std::chrono::steady_clock::now();
        const int progress_every = std::max(1, N_ITER / 10);
        const long ck = pars.checkpoint_every;
        auto is_event = [&](int t) {
            return t % pars.ANAL_EVERY == 0 || t % pars.dump_interval == 0 ||
                   t % progress_every == 0 || (ck > 0 && t % ck == 0) || t == N_ITER - 1;
        };
        int t = t0;
        while (t < N_ITER) {
            int te = t;
            while (te < N_ITER - 1 && !is_event(te))
                te++;
            state.step_rk4(te - t + 1); // steps t .. te
            t = te;
            if (t % pars.ANAL_EVERY == 0) {
                state.analyse((unsigned)t);
                state.surv_prob((unsigned)t);
            }
            if (t % pars.dump_interval == 0)
                state.dump_rho((unsigned)t);
            if (ck > 0 && t % ck == 0)
                state.write_state((unsigned)(t + 1));
            if (t % progress_every == 0)
                std::cout << " " << std::setprecision(4)
                          << (int)(t * 100.0 / (N_ITER - 1 > 0 ? N_ITER - 1 : 1)) << " %"
                          << std::endl;
            t++;
        }
        check(cosenu_sync(state.ctx), "cosenu_sync");

        std::cout << " 100 %" << std::endl << std::endl << "SIMULATION COMPLETED" << std::endl;
        auto end = std::chrono::steady_clock::now();
        std::chrono::duration<double> el = end - start;
        std::ofstream xtf("time.txt", std::ofstream::out | std::ofstream::trunc);
        xtf << el.count() << std::endl;
        std::cout << "elapsed time: " << el.count() << "s\n";
    } catch (const std::exception &e) {
        std::cout << "[ FAIL ]..." << e.what() << std::endl;
        return 1;
    }
    return 0;
}
