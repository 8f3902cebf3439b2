// bec_run: the driver surface of the reference program
// (programs/bec_wigner_3d.cuda/wigner.cpp:313-369) on top of the C ABI:
// YAML experiment parameters -> ground state -> ensemble initialisation and
// pi/2 pulse -> evolution with sampled observables -> YAML xy series in the
// schema of the reference's result files (data/gpu_sp_30k.yaml: xarray,
// yarray, xname, yname, name, description, source, format: xy, ymin, ymax).
//
// usage: bec_run experiment.yaml [output_prefix]
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "bec_api.h"

namespace {

// Flat "key: value" YAML subset: comments (#), blank lines, quoted or bare scalars.
std::map<std::string, std::string> read_yaml(const char *path, bool *ok)
{
    std::map<std::string, std::string> kv;
    *ok = false;
    FILE *f = fopen(path, "r");
    if (!f) return kv;
    char line[4096];
    while (fgets(line, sizeof(line), f)) {
        std::string s(line);
        size_t hash = s.find('#');
        if (hash != std::string::npos) s = s.substr(0, hash);
        size_t colon = s.find(':');
        if (colon == std::string::npos) continue;
        auto trim = [](std::string t) {
            size_t a = t.find_first_not_of(" \t\r\n"), b = t.find_last_not_of(" \t\r\n");
            if (a == std::string::npos) return std::string();
            t = t.substr(a, b - a + 1);
            if (t.size() >= 2 && (t[0] == '\'' || t[0] == '"') && t.back() == t[0]) t = t.substr(1, t.size() - 2);
            return t;
        };
        std::string k = trim(s.substr(0, colon)), v = trim(s.substr(colon + 1));
        if (!k.empty()) kv[k] = v;
    }
    fclose(f);
    *ok = true;
    return kv;
}

double num(const std::map<std::string, std::string> &kv, const char *k, double def)
{
    auto it = kv.find(k);
    return it == kv.end() || it->second.empty() ? def : strtod(it->second.c_str(), nullptr);
}

std::string str(const std::map<std::string, std::string> &kv, const char *k, const char *def)
{
    auto it = kv.find(k);
    return it == kv.end() ? std::string(def) : it->second;
}

bool flag(const std::map<std::string, std::string> &kv, const char *k, bool def)
{
    auto it = kv.find(k);
    if (it == kv.end()) return def;
    const std::string &v = it->second;
    return v == "1" || v == "true" || v == "True" || v == "yes" || v == "on";
}

std::string list(const std::vector<double> &v)
{
    std::string s = "[";
    char buf[64];
    for (size_t i = 0; i < v.size(); ++i) {
        snprintf(buf, sizeof(buf), "%.17g", v[i]);
        s += buf;
        if (i + 1 < v.size()) s += ", ";
    }
    return s + "]";
}

// The xy schema of the reference's data/*.yaml; `yerrors` (standard errors of
// the ensemble means) is an extra key written only when errors are given.
int write_series(const std::string &path, const std::vector<double> &x, const std::vector<double> &y,
                 const char *xname, const char *yname, const std::string &name, const std::string &desc,
                 double ymin, double ymax, const std::vector<double> *yerr = nullptr)
{
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return 1;
    fprintf(f, "xarray: '%s'\n", list(x).c_str());
    fprintf(f, "xname: %s\n", xname);
    fprintf(f, "yarray: '%s'\n", list(y).c_str());
    if (yerr) fprintf(f, "yerrors: '%s'\n", list(*yerr).c_str());
    fprintf(f, "ymax: %.17g\nymin: %.17g\n", ymax, ymin);
    fprintf(f, "yname: %s\n", yname);
    fprintf(f, "description: '%s'\n", desc.c_str());
    fprintf(f, "format: xy\n");
    fprintf(f, "name: '%s'\n", name.c_str());
    fprintf(f, "source: bec_run\n");
    fclose(f);
    return 0;
}

#define CHECK(call)                                                                                    \
    do {                                                                                               \
        int st_ = (call);                                                                              \
        if (st_ != BEC_OK) {                                                                           \
            fprintf(stderr, "%s failed: %s\n", #call, bec_strerror(st_));                              \
            return 1;                                                                                  \
        }                                                                                              \
    } while (0)

} // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s experiment.yaml [output_prefix]\n", argv[0]);
        return 2;
    }
    bool ok = false;
    auto kv = read_yaml(argv[1], &ok);
    if (!ok) {
        fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    const std::string prefix = argc > 2 ? argv[2] : str(kv, "output", "bec_run");

    bec_model_params m;
    bec_default_model_params(&m);
    m.a11 = num(kv, "a11", m.a11);
    m.a12 = num(kv, "a12", m.a12);
    m.a22 = num(kv, "a22", m.a22);
    m.m = num(kv, "m", m.m);
    m.fx = num(kv, "fx", m.fx);
    m.fy = num(kv, "fy", m.fy);
    m.fz = num(kv, "fz", m.fz);
    m.N = (int)num(kv, "N", m.N);
    m.detuning = num(kv, "detuning", m.detuning);
    m.gamma111 = num(kv, "gamma111", m.gamma111);
    m.gamma12 = num(kv, "gamma12", m.gamma12);
    m.gamma22 = num(kv, "gamma22", m.gamma22);
    m.Va = num(kv, "Va", m.Va);
    m.Vb = num(kv, "Vb", m.Vb);
    m.nvx = (int)num(kv, "nvx", m.nvx);
    m.nvy = (int)num(kv, "nvy", m.nvy);
    m.nvz = (int)num(kv, "nvz", m.nvz);
    m.itmax = (int)num(kv, "itmax", m.itmax);
    m.dt_steady = num(kv, "dt_steady", m.dt_steady);
    m.t_equilib = num(kv, "t_equilib", m.t_equilib);
    m.dt_evo = num(kv, "dt_evo", m.dt_evo);
    m.ne = (int)num(kv, "ne", m.ne);
    bec_calc_params c;
    CHECK(bec_fill_derived_params(&m, &c));
    // the run's length and sampling period are checked before any device work:
    // a period of 0 would never advance the loop below
    const double steps_in = num(kv, "steps", 15000), every_in = num(kv, "sample_every", 250);
    if (!(steps_in >= 0) || steps_in > 2e9 || steps_in != floor(steps_in)) {
        fprintf(stderr, "steps must be a non-negative integer (got %s)\n", str(kv, "steps", "").c_str());
        return 2;
    }
    if (!(every_in >= 1) || every_in > 2e9 || every_in != floor(every_in)) {
        fprintf(stderr, "sample_every must be a positive integer (got %s)\n", str(kv, "sample_every", "").c_str());
        return 2;
    }
    const int steps = (int)steps_in, every = (int)every_in;
    printf("mu(TF) = %f\n", c.mu);

    const int prec = str(kv, "precision", "fp32") == "fp64" ? BEC_FP64 : BEC_FP32;
    bec_ctx *ctx = nullptr;
    CHECK(bec_create(&c, prec, nullptr, &ctx));

    bec_gs_options o;
    bec_default_gs_options(&o);
    if (str(kv, "gs_method", "split_step") == "gradient") {
        o.method = BEC_GS_GRADIENT;
        o.rule = BEC_GS_RULE_RESIDUAL;
    }
    if (str(kv, "gs_rule", o.rule == BEC_GS_RULE_RESIDUAL ? "residual" : "reference") == "residual")
        o.rule = BEC_GS_RULE_RESIDUAL;
    o.tol = num(kv, "gs_tol", o.tol);
    o.max_iter = (int)num(kv, "gs_max_iter", o.max_iter);
    std::vector<double> gs(2 * (size_t)c.cells);
    bec_gs_report rep;
    int st = bec_calculate_steady_state(ctx, 0, nullptr, &o, gs.data(), &rep);
    if (st != BEC_OK && st != BEC_ERR_NOT_CONVERGED) {
        fprintf(stderr, "ground state failed: %s\n", bec_strerror(st));
        return 1;
    }
    printf("Steady state: %d iterations, %.3f ms, E = %f, mu = %f, residual %.3e\n", rep.iterations, rep.elapsed_ms,
           rep.E_ref, rep.mu_ref, rep.residual);

    const uint64_t seed = (uint64_t)num(kv, "seed", 0x5EED);
    // loss-induced Wigner noise (off by default, as in the reference's kernel);
    // switched on before the ensembles so that the equilibration steps carry it too
    if (flag(kv, "loss_noise", false))
        CHECK(bec_set_loss_noise(ctx, 1, (uint64_t)num(kv, "loss_noise_seed", (double)(seed + 1)), 0));
    CHECK(bec_init_evolution(ctx, 0, gs.data(), seed, flag(kv, "equilibrate", true) ? 1 : 0));

    std::vector<double> t, vis, visg, ntot, na, nb, phase, na_err, nb_err;
    bec_obs ob;
    CHECK(bec_observables(ctx, 0, &ob));
    auto record = [&](const bec_obs &o2) {
        t.push_back(o2.t * c.t_rho);
        vis.push_back(o2.visibility);
        visg.push_back(o2.visibility_grid);
        ntot.push_back(o2.N_total);
        na.push_back(o2.Na / c.ne);
        nb.push_back(o2.Nb / c.ne);
        na_err.push_back(o2.Na_stderr);
        nb_err.push_back(o2.Nb_stderr);
        phase.push_back(o2.phase);
    };
    record(ob);
    for (int done = 0; done < steps;) {
        const int n = std::min(every, steps - done);
        CHECK(bec_propagate(ctx, n, 0.0, 0, nullptr));
        done += n;
        CHECK(bec_observables(ctx, 0, &ob));
        record(ob);
        printf("%f %f %f\n", ob.t * c.t_rho * 1000.0, ob.N_total, ob.visibility_grid);
    }
    const std::string name = str(kv, "name", "bec_run");
    char desc[256];
    snprintf(desc, sizeof(desc), "MI355X, %s, %dx%dx%d, %g s timestep, %d trajectories", prec ? "fp64" : "fp32",
             c.nvx, c.nvy, c.nvz, m.dt_evo, c.ne);
    int w = 0;
    w |= write_series(prefix + "_visibility.yaml", t, visg, "Time, s", "Visibility", name, desc, 0.0, 1.0);
    w |= write_series(prefix + "_visibility_analytic.yaml", t, vis, "Time, s", "Visibility", name, desc, 0.0, 1.0);
    w |= write_series(prefix + "_N.yaml", t, ntot, "Time, s", "N", name, desc, 0.0, (double)c.N);
    w |= write_series(prefix + "_Na.yaml", t, na, "Time, s", "N", name, desc, 0.0, (double)c.N, &na_err);
    w |= write_series(prefix + "_Nb.yaml", t, nb, "Time, s", "N", name, desc, 0.0, (double)c.N, &nb_err);
    w |= write_series(prefix + "_phase.yaml", t, phase, "Time, s", "Phase", name, desc, -M_PI, M_PI);
    bec_destroy(ctx);
    if (w) {
        fprintf(stderr, "cannot write output series\n");
        return 1;
    }
    return 0;
}
