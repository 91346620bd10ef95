/*
 * cosenu_nuosc_binding.hpp -- the reference-side binding of COSEnu's
 * `class NuOsc` (src/nuosc.hpp:7-191) to the C ABI of cosenu_hip.h.
 *
 * Header-only C++ for the reference's own translation unit: include it after
 * the reference's headers (src/presets.hpp, structures.hpp, parser.hpp,
 * nuosc.hpp); it reads the preset macros and the members of NuOsc and Params
 * and calls nothing but the C ABI.  A maintainer who edits class NuOsc adds
 *
 *     cosenu_ctx *gpu = nullptr;           // member
 *     cosenu_params gpu_par;               // member, cosenu_binding::params_of(pars, mode)
 *     bool gpu_ahead = false;              // member: device state newer than v_stat
 *
 * and replaces the body of NuOsc::step_rk4 (nuosc.hpp:320-348) by
 * cosenu_binding::step(*this, gpu, gpu_par, gpu_ahead), and puts
 * cosenu_binding::fetch(*this, gpu, gpu_ahead) at the top of the host readers
 * of v_stat (analyse, surv_prob, dump_rho, write_state).
 * oracle/ref_binding.cpp applies exactly this to the reference's unmodified
 * sources (as a subclass) and drives the reference's own main loop.
 *
 * Failures print the library's message in the reference's "[ FAIL ]..." style
 * and exit, as the reference does on its own errors (parser.hpp:107-113).
 */
#ifndef COSENU_NUOSC_BINDING_HPP
#define COSENU_NUOSC_BINDING_HPP

#include <cstdlib>
#include <iostream>

#include "cosenu_hip.h"

namespace cosenu_binding {

/* The compile-time presets (src/presets.hpp, written by manage.py:160-186) as
 * the run-time fields of cosenu_params. */
inline int preset_scheme()
{
#if defined(FD)
    return COSENU_FD;
#else
    return COSENU_FV;
#endif
}

inline int preset_bc()
{
#if defined(OPEN_BC)
    return COSENU_OPEN;
#else
    return COSENU_PERIODIC;
#endif
}

inline int preset_flags()
{
    int f = 0;
#if defined(VAC_OSC_ON)
    f |= COSENU_VAC;
#endif
#if defined(MAT_OSC_ON)
    f |= COSENU_MAT;
#endif
#if defined(COLL_OSC_ON)
    f |= COSENU_COLL;
#endif
#if defined(ADVEC_OFF)
    f |= COSENU_ADVEC_OFF;
#endif
#if defined(FD) && !defined(KO_ORD_3)
    f |= COSENU_KO5;
#endif
    return f;
}

inline void check(int rc)
{
    if (rc) {
        std::cout << "[ FAIL ]...cosenu: " << cosenu_last_error() << std::endl;
        std::exit(1);
    }
}

/* Run parameters from the job.config values (parser.hpp:46-97): omega and
 * theta (degrees) come from Params because NuOsc keeps only their products
 * ct, st (set_vac_pars, nuosc.hpp:132-138); ko is NuOsc's 0.1
 * (nuosc.hpp:60).  The vacuum and collective values only matter when their
 * presets are on, exactly as set_vac_pars / set_collective_pars are only
 * called then (main.cpp:120-127). */
template <class P>
cosenu_params params_of(const P &pars, int mode, int device = 0)
{
    cosenu_params p;
    p.scheme = preset_scheme();
    p.bc = preset_bc();
    p.flags = preset_flags();
    p.nz = pars.nz;
    p.nvz = pars.nvz;
    p.gz = pars.gz;
    p.z0 = pars.z0;
    p.z1 = pars.z1;
    p.v0 = pars.v0;
    p.v1 = pars.v1;
    p.CFL = pars.CFL;
    p.pmo = pars.pmo;
    p.omega = pars.omega;
    p.theta = pars.theta;
    p.mu = pars.mu;
    p.ko = 1.0e-1;
    p.mode = mode;
    p.device = device;
    return p;
}

/* The 8 planes of a FieldVar (structures.hpp:5-37) in the ABI's field order.
 * (Template parameters avoid the preset macro names FV and FD.) */
template <class FieldVarT>
void planes_of(const FieldVarT *v, double *out[8])
{
    out[0] = v->ee; out[1] = v->xx; out[2] = v->ex_re; out[3] = v->ex_im;
    out[4] = v->bee; out[5] = v->bxx; out[6] = v->bex_re; out[7] = v->bex_im;
}

/* Create the device state from the host's profiles and v_stat (the rows of a
 * FieldVar are nz + 2 gz long with the interior at offset gz,
 * nuosc.hpp:145-148). */
template <class Osc>
void open(Osc &s, cosenu_ctx *&gpu, const cosenu_params &par)
{
    check(cosenu_create(&par, &gpu));
    const ptrdiff_t R = (ptrdiff_t)s.nz + 2 * s.gz;
    check(cosenu_set_profile(gpu, s.G0->G, s.G0->bG, R, s.gz));
    double *p[8];
    planes_of(s.v_stat, p);
    check(cosenu_upload_planes(gpu, COSENU_STATE, p, R, s.gz));
}

/* NuOsc::step_rk4 (nuosc.hpp:320-348) on the device: the first call uploads
 * the host state (after initialize or read_state), later calls only enqueue;
 * phy_time advances as the reference's does. */
template <class Osc>
void step(Osc &s, cosenu_ctx *&gpu, const cosenu_params &par, bool &ahead)
{
    if (!gpu)
        open(s, gpu, par);
    check(cosenu_step(gpu, 1));
    ahead = true;
    s.phy_time += s.dt;
}

/* Bring v_stat on the host up to date before the host reads it. */
template <class Osc>
void fetch(Osc &s, cosenu_ctx *gpu, bool &ahead)
{
    if (!gpu || !ahead)
        return;
    double *p[8];
    planes_of(s.v_stat, p);
    check(cosenu_download_planes(gpu, COSENU_STATE, p, (ptrdiff_t)s.nz + 2 * s.gz, s.gz));
    ahead = false;
}

inline void close(cosenu_ctx *&gpu)
{
    if (gpu)
        cosenu_destroy(gpu);
    gpu = nullptr;
}

} // namespace cosenu_binding

#endif /* COSENU_NUOSC_BINDING_HPP */
