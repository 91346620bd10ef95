// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Builds the reference's own header-only solver (src/*.hpp under the
// reference checkout) into a shared library with a C entry point per
// operation, so that tests can compare the oracle restatement bit for bit
// with the reference code.  Nothing of the reference is copied: the headers
// are included from where they lie (oracle/Makefile passes -I to them).
//
// The preset macros (FV/FD, PERIODIC_BC/OPEN_BC, VAC_OSC_ON, MAT_OSC_ON,
// COLL_OSC_ON, ADVEC_OFF, KO_ORD_3, ADVEC_CENTER_FD) come from the command
// line, exactly the set that manage.py:160-186 writes into presets.hpp.
//
// State exchange format: state[f][i][j] (8 fields, nvz, nz), no ghosts.

#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <vector>
#include <chrono>

using std::cin;
using std::cout;
using std::endl;
using std::string;

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

namespace {

double **planes(FieldVar *v)
{
    static thread_local double *p[8];
    p[0] = v->ee; p[1] = v->xx; p[2] = v->ex_re; p[3] = v->ex_im;
    p[4] = v->bee; p[5] = v->bxx; p[6] = v->bex_re; p[7] = v->bex_im;
    return p;
}

void load(NuOsc &s, FieldVar *dst, const double *state)
{
    double **p = planes(dst);
    const size_t P = (size_t)s.nvz * s.nz;
    for (int f = 0; f < 8; f++)
        for (int i = 0; i < s.nvz; i++)
            for (int j = 0; j < s.nz; j++)
                p[f][s.idx(i, j)] = state[f * P + (size_t)i * s.nz + j];
}

void store(NuOsc &s, FieldVar *src, double *state)
{
    double **p = planes(src);
    const size_t P = (size_t)s.nvz * s.nz;
    for (int f = 0; f < 8; f++)
        for (int i = 0; i < s.nvz; i++)
            for (int j = 0; j < s.nz; j++)
                state[f * P + (size_t)i * s.nz + j] = p[f][s.idx(i, j)];
}

void load_profile(NuOsc &s, const double *G, const double *bG)
{
    for (int i = 0; i < s.nvz; i++)
        for (int j = 0; j < s.nz; j++) {
            s.G0->G[s.idx(i, j)] = G[(size_t)i * s.nz + j];
            s.G0->bG[s.idx(i, j)] = bG[(size_t)i * s.nz + j];
        }
}

// Parameter set-up done by main.cpp:111-127.
void setup(NuOsc &s, const Params &p)
{
#ifdef VAC_OSC_ON
    s.set_vac_pars(p.pmo, p.omega, p.theta);
#endif
#ifdef COLL_OSC_ON
    s.set_collective_pars(p.mu);
#endif
    (void)p;
}

} // namespace

extern "C" {

// Bit mask of the presets this library was built with.
int ref_presets(void)
{
    int m = 0;
#ifdef FD
    m |= 1;
#endif
#ifdef OPEN_BC
    m |= 2;
#endif
#ifdef VAC_OSC_ON
    m |= 4;
#endif
#ifdef MAT_OSC_ON
    m |= 8;
#endif
#ifdef COLL_OSC_ON
    m |= 16;
#endif
#ifdef ADVEC_OFF
    m |= 32;
#endif
#ifdef KO_ORD_3
    m |= 64;
#endif
    return m;
}

// out = calRHS(in) after the ghost fill of step_rk4 (nuosc.hpp:323-324).
int ref_rhs(const char *conf, const double *in, double *out)
{
    Params p{std::string(conf)};
    NuOsc s(p);
    setup(s, p);
    load(s, s.v_stat, in);
    s.updateBufferZone(s.v_stat);
    s.calRHS(s.v_rhs, s.v_stat);
    store(s, s.v_rhs, out);
    return 0;
}

// nsteps calls of step_rk4 (main.cpp:175-177) from the given state.
int ref_step(const char *conf, double *state, int nsteps)
{
    Params p{std::string(conf)};
    NuOsc s(p);
    setup(s, p);
    load(s, s.v_stat, state);
    for (int t = 0; t < nsteps; t++)
        s.step_rk4();
    store(s, s.v_stat, state);
    return 0;
}

// NuOsc::initialize (writes G0.bin into the working directory, as the
// reference does).
int ref_initialize(const char *conf, double *G, double *bG, double *state)
{
    Params p{std::string(conf)};
    NuOsc s(p);
    setup(s, p);
    s.initialize();
    store(s, s.v_stat, state);
    for (int i = 0; i < s.nvz; i++)
        for (int j = 0; j < s.nz; j++) {
            G[(size_t)i * s.nz + j] = s.G0->G[s.idx(i, j)];
            bG[(size_t)i * s.nz + j] = s.G0->bG[s.idx(i, j)];
        }
    return 0;
}

// cal_pol + cal_Mn + dcon + surv_prob as main.cpp:157-186 calls them.  dcon
// and surv_prob append to conserved_quantities.dat / survival_probability.dat
// in the working directory (t = 0 truncates).  m0[4] = {e1, e2, e3, norm}.
int ref_analyse(const char *conf, const double *G, const double *bG, const double *state0,
                const double *state, int t, double *m0)
{
    Params p{std::string(conf)};
    NuOsc s(p);
    setup(s, p);
    load_profile(s, G, bG);
    FieldVar *v0 = new FieldVar(s.size);
    load(s, v0, state0);
    load(s, s.v_stat, state);
    Pol *P0 = new Pol(s.size);
    s.cal_pol(v0, P0);
    Pol *P = new Pol(s.size);
    M *M_0 = new M(0);
    s.cal_pol(s.v_stat, P);
    s.cal_Mn(M_0, P, 0);
    m0[0] = M_0->e1;
    m0[1] = M_0->e2;
    m0[2] = M_0->e3;
    m0[3] = M_0->norm;
    s.analyse(s.v_stat, P0, 0, t);
    s.surv_prob(s.v_stat, v0, t);
    delete P;
    delete M_0;
    delete P0;
    delete v0;
    return 0;
}

// write_state0 / read_state / read_G0 round trip through the reference's
// binary file code (snaps.hpp:43-183), in the working directory.
int ref_write_state0(const char *conf, const double *state)
{
    Params p{std::string(conf)};
    NuOsc s(p);
    FieldVar *v0 = new FieldVar(s.size);
    load(s, v0, state);
    s.write_state0(v0);
    delete v0;
    return 0;
}

int ref_read_state(const char *conf, double *state)
{
    Params p{std::string(conf)};
    NuOsc s(p);
    int t = s.read_state();
    store(s, s.v_stat, state);
    return t;
}

int ref_dump_rho(const char *conf, const double *state, int t)
{
    Params p{std::string(conf)};
    NuOsc s(p);
    load(s, s.v_stat, state);
    s.dump_rho(s.v_stat, (uint)t);
    return 0;
}

} // extern "C"
