// ref_binding.hpp -- TEST INFRASTRUCTURE ONLY.
//
// Prelude for building the reference's own main program (src/main.cpp, left
// as it lies under the reference checkout) with NuOsc::step_rk4 on the
// MI355X library through include/cosenu_nuosc_binding.hpp:
//
//     g++ -O3 -std=c++0x -fopenmp -D__PRESETS__ <presets> -I<ref>/src -I<repo>/include \
//         -include oracle/ref_binding.hpp <ref>/src/main.cpp -lcosenu_hip
//
// (oracle/Makefile, target `binding`).  The reference's headers are included
// here with their class renamed NuOscReference; NuOsc is then the subclass
// below, which applies the binding a maintainer would write into the class
// itself: step_rk4 on the device, v_stat fetched before the host reads it.
// main.cpp's own includes of the headers are then no-ops (include guards),
// and its `NuOsc state(pars)` is this class.  Nothing of the reference is
// copied or modified.
//
// COSENU_BINDING_MODE=strict selects the library's STRICT mode (the
// reference's arithmetic bit for bit); anything else FAST.
#pragma once

#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>
#include <chrono>

using std::cin;
using std::cout;
using std::endl;
using std::string;

#define NuOsc NuOscReference
#include "structures.hpp"
#include "miscellaneous_funcs.hpp"
#include "parser.hpp"
#include "nuosc.hpp"
#if defined(FD)
#include "rhs_fd.hpp"
#elif defined(FV)
#include "rhs_fv.hpp"
#endif
#include "initialize.hpp"
#include "snaps.hpp"
#include "analysis.hpp"
#undef NuOsc

#include "cosenu_nuosc_binding.hpp"

inline int cosenu_binding_mode()
{
    const char *m = std::getenv("COSENU_BINDING_MODE");
    return (m && std::strcmp(m, "strict") == 0) ? COSENU_STRICT : COSENU_FAST;
}

class NuOsc : public NuOscReference
{
public:
    cosenu_ctx *gpu = nullptr;
    cosenu_params gpu_par;
    bool gpu_ahead = false;

    NuOsc(Params pars)
        : NuOscReference(pars), gpu_par(cosenu_binding::params_of(pars, cosenu_binding_mode()))
    {
    }
    ~NuOsc() { cosenu_binding::close(gpu); }

    void step_rk4() { cosenu_binding::step(*this, gpu, gpu_par, gpu_ahead); }

    void analyse(const FieldVar *v, const Pol *P0, uint n, uint t)
    {
        cosenu_binding::fetch(*this, gpu, gpu_ahead);
        NuOscReference::analyse(v, P0, n, t);
    }
    void surv_prob(const FieldVar *v, const FieldVar *v0, const uint t)
    {
        cosenu_binding::fetch(*this, gpu, gpu_ahead);
        NuOscReference::surv_prob(v, v0, t);
    }
    void dump_rho(const FieldVar *v, const uint t)
    {
        cosenu_binding::fetch(*this, gpu, gpu_ahead);
        NuOscReference::dump_rho(v, t);
    }
    void write_state(unsigned int t)
    {
        cosenu_binding::fetch(*this, gpu, gpu_ahead);
        NuOscReference::write_state(t);
    }
};
