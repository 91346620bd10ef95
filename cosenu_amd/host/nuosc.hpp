// Host-side state object: the role of the reference's `class NuOsc`
// (src/nuosc.hpp:7-191) with its I/O (src/snaps.hpp), initial condition
// (src/initialize.hpp) and analysis writers (src/analysis.hpp), all driving
// the device through the C ABI of include/cosenu_hip.h.  File formats are the
// reference's byte for byte:
//   state.bin   int32 iteration, then for i < nvz, j < nz: 8 doubles
//               (ee, xx, ex_re, ex_im, bee, bxx, bex_re, bex_im)   snaps.hpp:6-39
//   state0.bin  the same body without the header                   snaps.hpp:43-71
//   G0.bin      for i < nvz, j < nz: (G, bG)                        initialize.hpp:40-41
//   rho_<t>.dat text, j outer, i inner, 16 significant digits       snaps.hpp:187-218
//   conserved_quantities.dat, survival_probability.dat              analysis.hpp:71-185
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cosenu_hip.h"
#include "params.hpp"

namespace cosenu_host {

inline void check(int rc, const char *what)
{
    if (rc != 0)
        throw std::runtime_error(std::string(what) + ": " + cosenu_last_error());
}

inline bool file_exists(const std::string &path)
{
    std::ifstream f(path.c_str(), std::ios::in);
    return (bool)f;
}

// Normalised Gaussian angular profile on [-1, 1] (miscellaneous_funcs.hpp:57-62).
inline double g_profile(double v, double v0, double sigma)
{
    double ex = (v - v0) * (v - v0) / (2.0 * sigma * sigma);
    double N = sigma * sqrt(M_PI / 2.0) *
               (erf((1.0 + v0) / sigma / sqrt(2.0)) + erf((1.0 - v0) / sigma / sqrt(2.0)));
    return exp(-ex) / N;
}

class NuOsc {
public:
    Params P;
    cosenu_ctx *ctx = nullptr;
    int nz, nvz;
    double dz, dv, dt, phy_time = 0.0;
    std::vector<double> vz, Z;
    std::vector<double> G, bG; // [nvz * nz], the profiles the analysis divides by
    size_t ncell;

    explicit NuOsc(const Params &p, int device = 0) : P(p), nz(p.nz), nvz(p.nvz)
    {
        cosenu_params cp = p.c(device);
        check(cosenu_create(&cp, &ctx), "cosenu_create");
        cosenu_info info;
        check(cosenu_get_info(ctx, &info), "cosenu_get_info");
        dz = info.dz;
        dv = info.dv;
        dt = info.dt;
        ncell = (size_t)nz * nvz;
        vz.resize(nvz);
        Z.resize(nz);
        for (int i = 0; i < nvz; i++)
            vz[i] = p.v0 + (i + 0.5) * dv;
        for (int j = 0; j < nz; j++)
            Z[j] = p.z0 + (j + 0.5) * dz;
        G.assign(ncell, 0.0);
        bG.assign(ncell, 0.0);
        banner();
    }
    ~NuOsc()
    {
        if (ctx)
            cosenu_destroy(ctx);
    }
    NuOsc(const NuOsc &) = delete;
    NuOsc &operator=(const NuOsc &) = delete;

    void banner() const
    {
        std::string bar(50, '#'), dots(50, '.');
        std::cout << bar << std::endl
                  << std::setw(35) << "Simulation state created" << std::endl
                  << dots << std::endl
                  << "\tSCHEME: " << P.scheme_name << std::endl
                  << "\tz : (" << P.z0 << ", " << P.z1 << ")" << std::endl
                  << "\tvz: (" << vz[0] << ", " << vz[nvz - 1] << ")" << std::endl
                  << "\t(Nvz, Nz): (" << nvz << ", " << nz << ")" << std::endl
                  << "\tgz: " << P.gz << std::endl
                  << "\tCFL: " << P.CFL << "\n\tperturbation_size: " << P.perturbation_size
                  << "\n\tmode: " << (P.mode == COSENU_FAST ? "fast" : "strict") << std::endl
                  << bar << std::endl;
    }

    // NuOsc::initialize (initialize.hpp:6-47): G = g(v;1,0.6), bG = 0.9 g(v;1,0.6),
    // ee/xx = G(1 +- sqrt(1-eps^2))/2, ex_re = G eps/2, eps = amp exp(-z^2/50).
    // Also writes G0.bin, uploads v_stat, and copies it to v_stat0 (copy_state,
    // main.cpp:139).
    void initialize(double signu = 0.6, double sigbnu = 0.6, double alpha = 0.90)
    {
        std::vector<double> buf(ncell * 8);
        std::ofstream g_file("G0.bin", std::ofstream::out | std::ofstream::binary);
        if (!g_file)
            std::cout << "Unable to open G0.bin file from NuOsc::initialise."
                      << "Will not be storing initial angular profiles.\n";
        for (int i = 0; i < nvz; i++)
            for (int j = 0; j < nz; j++) {
                size_t ij = (size_t)i * nz + j;
                G[ij] = g_profile(vz[i], 1.0, signu);
                bG[ij] = alpha * g_profile(vz[i], 1.0, sigbnu);
                double e = P.perturbation_size * exp(-(Z[j] - 0.0) * (Z[j] - 0.0) / 50.0);
                double e_ = sqrt(1.0 - e * e);
                double *c = &buf[ij * 8];
                c[0] = 0.5 * G[ij] * (1.0 + e_);
                c[1] = 0.5 * G[ij] * (1.0 - e_);
                c[2] = 0.5 * G[ij] * (0.0 + e);
                c[3] = -0.0;
                c[4] = 0.5 * bG[ij] * (1.0 + e_);
                c[5] = 0.5 * bG[ij] * (1.0 - e_);
                c[6] = 0.5 * bG[ij] * (0.0 + e);
                c[7] = 0.0;
                g_file.write((const char *)&G[ij], sizeof(double));
                g_file.write((const char *)&bG[ij], sizeof(double));
            }
        g_file.close();
        check(cosenu_set_profile(ctx, G.data(), bG.data(), nz, 0), "cosenu_set_profile");
        check(cosenu_upload_interleaved(ctx, COSENU_STATE, buf.data()), "upload state");
        check(cosenu_upload_interleaved(ctx, COSENU_STATE0, buf.data()), "upload state0");
        std::cout << "Simulation state initialized." << std::endl;
    }

    // snaps.hpp:43-71
    void write_state0() const
    {
        std::vector<double> buf(ncell * 8);
        check(cosenu_download_interleaved(ctx, COSENU_STATE0, buf.data()), "download state0");
        std::ofstream f("state0.bin", std::ios::out | std::ios::binary);
        if (!f) {
            std::cout << "Failed call to open state0.bin in NuOsc::write_state0. \n";
            return;
        }
        f.write((const char *)buf.data(), buf.size() * sizeof(double));
    }

    // snaps.hpp:6-39: header int32 t, then the body; written to state<t>.bin
    // and renamed over state.bin, so a crash never leaves a torn checkpoint.
    void write_state(unsigned int t) const
    {
        std::vector<double> buf(ncell * 8);
        check(cosenu_download_interleaved(ctx, COSENU_STATE, buf.data()), "download state");
        std::string tmp = "state" + std::to_string(t) + ".bin";
        {
            std::ofstream f(tmp.c_str(), std::ios::out | std::ios::binary);
            if (!f) {
                std::cout << "Failed call to open " << tmp << " in NuOsc::write_state. \n";
                return;
            }
            int ti = (int)t;
            f.write((const char *)&ti, sizeof(int));
            f.write((const char *)buf.data(), buf.size() * sizeof(double));
            if (!f)
                throw std::runtime_error("write_state: short write to " + tmp);
        }
        if (std::rename(tmp.c_str(), "state.bin") != 0)
            throw std::runtime_error(std::string("write_state: rename failed: ") + strerror(errno));
    }

    static void read_exact(std::ifstream &f, void *dst, size_t bytes, const std::string &name)
    {
        f.read((char *)dst, (std::streamsize)bytes);
        if ((size_t)f.gcount() != bytes)
            throw std::runtime_error(name + " is truncated: expected " + std::to_string(bytes) +
                                     " bytes, got " + std::to_string(f.gcount()));
    }

    // snaps.hpp:75-124: state.bin if present (returns its iteration), else
    // state0.bin with t = 1.
    int read_state()
    {
        std::string name = "state.bin";
        bool zero = false;
        if (!file_exists(name)) {
            std::cout << "Unable to find " << name << "\n";
            name = "state0.bin";
            zero = true;
        }
        std::ifstream f(name.c_str(), std::ios::in | std::ios::binary);
        if (!f)
            throw std::runtime_error(name + " did not open");
        int t = 1;
        if (!zero)
            read_exact(f, &t, sizeof(int), name);
        std::cout << "Loading from " << name << ". \n";
        std::vector<double> buf(ncell * 8);
        read_exact(f, buf.data(), buf.size() * sizeof(double), name);
        check(cosenu_upload_interleaved(ctx, COSENU_STATE, buf.data()), "upload state");
        return t;
    }

    // snaps.hpp:128-157
    void read_state0()
    {
        std::ifstream f("state0.bin", std::ios::in | std::ios::binary);
        if (!f)
            throw std::runtime_error("state0.bin does not exist");
        std::vector<double> buf(ncell * 8);
        read_exact(f, buf.data(), buf.size() * sizeof(double), "state0.bin");
        check(cosenu_upload_interleaved(ctx, COSENU_STATE0, buf.data()), "upload state0");
    }

    // snaps.hpp:161-183
    void read_G0()
    {
        std::ifstream f("G0.bin", std::ios::in | std::ios::binary);
        if (!f)
            throw std::runtime_error("G0.bin does not exist");
        std::vector<double> pairs(ncell * 2);
        read_exact(f, pairs.data(), pairs.size() * sizeof(double), "G0.bin");
        for (size_t k = 0; k < ncell; k++) {
            G[k] = pairs[2 * k];
            bG[k] = pairs[2 * k + 1];
        }
        check(cosenu_set_profile(ctx, G.data(), bG.data(), nz, 0), "cosenu_set_profile");
    }

    void step_rk4(int n = 1)
    {
        check(cosenu_step(ctx, n), "cosenu_step");
        phy_time += n * dt;
    }

    // analyse (analysis.hpp:3-25) -> conserved_quantities.dat (dcon format,
    // analysis.hpp:79-129); the reductions run on the device.
    void analyse(unsigned int t) const
    {
        cosenu_diag d;
        check(cosenu_diagnostics(ctx, &d), "cosenu_diagnostics");
        const char *name = "conserved_quantities.dat";
        std::ofstream f;
        f.open(name, t == 0 ? std::ofstream::out | std::ofstream::trunc
                            : std::ofstream::out | std::ofstream::app);
        if (!f) {
            std::cout << "Unable to open " << name << std::endl;
            return;
        }
        if (t == 0)
            f << "# [time, dP_max(t), <dP>(t), <dbP(t), <dP>(t), |M0|]\n";
        f << t << "\t" << std::fixed << std::setprecision(20) << d.dP_max << "\t" << d.dP_avg
          << "\t" << d.dbP_avg << "\t" << d.M0_norm << std::endl;
        if (d.nonfinite != 0)
            std::cout << "[ WARN ]...state holds " << d.nonfinite << " non-finite values at t = "
                      << t << std::endl;
    }

    // surv_prob (analysis.hpp:134-185) -> survival_probability.dat
    void surv_prob(unsigned int t) const
    {
        cosenu_diag d;
        check(cosenu_diagnostics(ctx, &d), "cosenu_diagnostics");
        const char *name = "survival_probability.dat";
        std::ofstream f;
        if (t == 0) {
            f.open(name, std::ofstream::out | std::ofstream::trunc);
            if (!f) {
                std::cout << "Unable to open " << name << std::endl;
                return;
            }
            f << "# [time, <Pee>, <Pbee>]" << std::endl;
        } else {
            f.open(name, std::ofstream::out | std::ofstream::app);
        }
        if (!f) {
            std::cout << "Unable to open " << name << std::endl;
            return;
        }
        f << t << "\t" << std::scientific << d.Pee << "\t" << d.Pbee << std::endl;
    }

    // dump_rho (snaps.hpp:187-218).  The stream's precision becomes 16 after
    // the first record's z and v are written, exactly as in the reference.
    void dump_rho(unsigned int t) const
    {
        std::vector<double> planes(ncell * 8);
        double *pp[8];
        for (int f = 0; f < 8; f++)
            pp[f] = &planes[(size_t)f * ncell];
        check(cosenu_download_planes(ctx, COSENU_STATE, pp, nz, 0), "download planes");
        std::string name = "rho_" + std::to_string(t) + ".dat";
        std::ofstream o(name.c_str(), std::ofstream::out | std::ofstream::trunc);
        if (!o) {
            std::cout << "Unable to open " << name << "\n";
            return;
        }
        for (int j = 0; j < nz; j++) {
            for (int i = 0; i < nvz; i++) {
                size_t ij = (size_t)i * nz + j;
                o << Z[j] << "\t" << vz[i] << "\t" << std::setprecision(16) << pp[0][ij] << "\t"
                  << pp[1][ij] << "\t" << pp[2][ij] << "\t" << pp[3][ij] << "\t" << pp[4][ij]
                  << "\t" << pp[5][ij] << "\t" << pp[6][ij] << "\t" << pp[7][ij] << "\n";
            }
            o << "\n";
        }
    }
};

} // namespace cosenu_host
