// Run configuration: the reference's job.config reader (src/parser.hpp:46-256)
// with the compile-time presets (src/presets.hpp, written by manage.py:160-186)
// turned into optional runtime keys.
//
// Same syntax as the reference: one `key : value` per line; unknown keys are
// reported and ignored (parser.hpp:211-215); required keys are
//   scheme nz nvz z0 z1 v0 v1 CFL gz N_ITER ANAL_EVERY
// plus pmo omega theta when the vacuum term is on and mu when the collective
// term is on (parser.hpp:220-243).  Unlike the reference, a missing key is an
// error returned to the caller (the reference exits with status 0).
//
// Runtime preset keys (defaults = the reference's checked-in presets.hpp:
// PERIODIC_BC, FV, VAC_OSC_ON, COLL_OSC_ON):
//   boundary : periodic | open
//   vac_osc_on / mat_osc_on / collective_osc_on / advection_off : True | False
//   ko_order : 3 | 5          (FD Kreiss-Oliger order; 3 = KO_ORD_3)
//   mode : strict | fast      (strict = bit-identical to the reference CPU build)
//   checkpoint_every : N      (write state.bin every N steps; 0 = never)
#pragma once

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "cosenu_hip.h"

namespace cosenu_host {

inline std::string trim(const std::string &s)
{
    size_t b = s.find_first_not_of(" \t\r");
    if (b == std::string::npos)
        return "";
    size_t e = s.find_last_not_of(" \t\r");
    return s.substr(b, e - b + 1);
}

inline bool parse_bool(const std::string &v)
{
    return v == "True" || v == "true" || v == "1" || v == "yes" || v == "on";
}

struct Params {
    std::string scheme_name;
    int N_ITER = 0, ANAL_EVERY = 0, nz = 0, nvz = 0, gz = 0;
    double z0 = 0, z1 = 0, v0 = 0, v1 = 0, CFL = 0;
    double pmo = 0, omega = 0, theta = 0, mu = 1.0, perturbation_size = 0;
    int n_dump = 0, dump_interval = 0;
    // runtime presets
    int scheme = COSENU_FV, bc = COSENU_PERIODIC, mode = COSENU_STRICT;
    bool vac = true, mat = false, coll = true, advec_off = false;
    int ko_order = 3;
    long checkpoint_every = 0;

    int flags() const
    {
        return (vac ? COSENU_VAC : 0) | (mat ? COSENU_MAT : 0) | (coll ? COSENU_COLL : 0) |
               (advec_off ? COSENU_ADVEC_OFF : 0) |
               ((scheme == COSENU_FD && ko_order == 5) ? COSENU_KO5 : 0);
    }

    cosenu_params c(int device = 0) const
    {
        cosenu_params p{};
        p.scheme = scheme;
        p.bc = bc;
        p.flags = flags();
        p.nz = nz;
        p.nvz = nvz;
        p.gz = gz;
        p.z0 = z0;
        p.z1 = z1;
        p.v0 = v0;
        p.v1 = v1;
        p.CFL = CFL;
        p.pmo = pmo;
        p.omega = omega;
        p.theta = theta;
        p.mu = mu;
        p.ko = 1.0e-1; // nuosc.hpp:60
        p.mode = mode;
        p.device = device;
        return p;
    }

    // Returns "" on success, else the reason.
    std::string load(const std::string &path)
    {
        std::ifstream in(path.c_str());
        if (!in)
            return "[ FAIL ]...Unable to open " + path;
        bool has[11] = {false};
        bool has_pmo = false, has_omega = false, has_theta = false, has_mu = false;
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#')
                continue;
            size_t pos = line.find(':');
            if (pos == std::string::npos)
                continue;
            std::string key = trim(line.substr(0, pos));
            std::string val = trim(line.substr(pos + 1));
            std::istringstream vs(val);
            if (key == "scheme") { vs >> scheme_name; has[0] = true;
                scheme = (scheme_name == "FD" || scheme_name == "fd") ? COSENU_FD : COSENU_FV; }
            else if (key == "nz") { vs >> nz; has[1] = true; }
            else if (key == "nvz") { vs >> nvz; has[2] = true; }
            else if (key == "z0") { vs >> z0; has[3] = true; }
            else if (key == "z1") { vs >> z1; has[4] = true; }
            else if (key == "v0") { vs >> v0; has[5] = true; }
            else if (key == "v1") { vs >> v1; has[6] = true; }
            else if (key == "CFL") { vs >> CFL; has[7] = true; }
            else if (key == "gz") { vs >> gz; has[8] = true; }
            else if (key == "N_ITER") { vs >> N_ITER; has[9] = true; }
            else if (key == "ANAL_EVERY") { vs >> ANAL_EVERY; has[10] = true; }
            else if (key == "pmo") { vs >> pmo; has_pmo = true; }
            else if (key == "omega") { vs >> omega; has_omega = true; }
            else if (key == "theta") { vs >> theta; has_theta = true; }
            else if (key == "mu") { vs >> mu; has_mu = true; }
            else if (key == "perturbation_size") vs >> perturbation_size;
            else if (key == "n_dump_rho") vs >> n_dump;
            else if (key == "boundary") bc = (val == "open") ? COSENU_OPEN : COSENU_PERIODIC;
            else if (key == "vac_osc_on") vac = parse_bool(val);
            else if (key == "mat_osc_on") mat = parse_bool(val);
            else if (key == "collective_osc_on") coll = parse_bool(val);
            else if (key == "advection_off") advec_off = parse_bool(val);
            else if (key == "ko_order") vs >> ko_order;
            else if (key == "mode") mode = (val == "fast") ? COSENU_FAST : COSENU_STRICT;
            else if (key == "checkpoint_every") vs >> checkpoint_every;
            else
                std::cout << "Redundant or unknown key: " << key << std::endl;
        }
        for (bool h : has)
            if (!h)
                return "[ FAIL ]...Incomplete config file";
        if (vac && !(has_pmo && has_omega && has_theta))
            return "[ FAIL ]...Incomplete config file";
        if (coll && !has_mu)
            return "[ FAIL ]...Incomplete config file";
        // parser.hpp:244-255
        if (n_dump > 0) {
            dump_interval = (int)(N_ITER / n_dump);
            if (dump_interval < 1)
                dump_interval = 1;
        } else {
            dump_interval = N_ITER + 1;
        }
        if (ANAL_EVERY < 1)
            return "[ FAIL ]...ANAL_EVERY must be at least 1";
        return "";
    }
};

} // namespace cosenu_host
