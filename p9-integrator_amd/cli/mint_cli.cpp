// mint_cli.cpp — the `mint` command line tool on top of libmint.
//
//   mint -c config.json [-t steps] [-d device] [-o positions_file] [-i] [-p]
//        [-s checkpoint_out] [-r checkpoint_in] [-n] [-a] [-g gpu]
//
// Same flags and JSON schema as the reference driver (README.md:10-19;
// args_parse src/simutils.cuh:138-175; sim_from_config_file
// src/simutils.cuh:232-301), with these documented differences:
//   -t is an exact step count (default 1, as README.md:18 says); the
//      reference counts 1e6-step launches (src/mercurius.cu:179).
//   -d D, as in the reference, selects the device and the sweeps [4 D, 4 D + 4)
//      of the file (SWEEPS_PER_GPU, src/constants.cuh:12; the slice of
//      src/simutils.cuh:288-290), each its own system, and names the output
//      files output_<D> and elements_<D>.txt; a slice past the end of the file
//      is an error (the reference reads past its sweep array there), a short
//      last slice runs the sweeps it has.  -a (addition): every sweep of the
//      file instead, as one batch (the 4096-system sweep).  -g G (addition):
//      run on device G instead of D, so that one GPU can reproduce the files
//      of any -d.  A sweep replaces body 0 including its mass (the reference
//      drops the sweep mass, src/mercurius.cu:126-164, and forces
//      mean_anomaly to 0, src/simutils.cuh:299).
//   optional keys: "dt" (default 0.8219, src/mercurius.cu:43), "n_hill",
//      "bs_tol", "softening2", "central_mass", "form" (plain|flat|batched),
//      "checkpoint_every" (steps between output checkpoints; default: end only),
//      per-body "radius" [AU] and top-level "density" [g/cm^3] (physical
//      radii, recorded and checked for overlaps, never merged; a body without
//      "radius" gets (3 m / (4 pi density))^(1/3) when "density" is given).
//   systems of more than 8 massive bodies (up to 2048) run when they carry no
//      test particles (the all-massive embryo configuration).
// -s / -r (additions): write a bit-exact binary checkpoint at the end of the
// run / resume from one (the state replaces the configuration's initial state;
// -c still names the bodies for the output files).  -n (addition): read the
// configuration, print what would run, and exit without touching a GPU.
// Output files keep the reference formats: positions appended to output_<dev>
// (src/simutils.cuh:186-230) and elements in elements_<dev>.txt
// (src/mercurius.cu:210-248; src/simutils.cuh:177-184).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mint.h"
#include "json_min.h"

namespace {

struct Body {
  std::string name;
  double mass, a, e, inc, node, argp, manom;
  double radius;  // -1: not given
};

Body body_from(const jsonmin::Value &v, const std::string &dflt_name) {
  Body b;
  b.name = v.string_or("name", dflt_name);
  b.mass = v.at("mass").number();
  b.a = v.at("semi_major_axis").number();
  b.e = v.at("eccentricity").number();
  b.inc = v.at("inclination").number();
  b.node = v.at("longitude_of_ascending_node").number();
  b.argp = v.at("argument_of_perihelion").number();
  b.manom = v.number_or("mean_anomaly", 0.0);
  b.radius = v.number_or("radius", -1.0);
  return b;
}

// sweeps each device integrates (SWEEPS_PER_GPU, src/constants.cuh:12)
constexpr int SWEEPS_PER_DEVICE = 4;

struct Args {
  std::string config, output, save, resume;
  long long steps = 1;
  int device = 0;
  int gpu = -1;  // -g: device to run on (default: -d)
  bool info = false, print_positions = false, dry = false, all_sweeps = false;
};

int usage() {
  std::cerr << "usage: mint -c config.json [-t steps] [-d device] [-o positions_file] [-i] [-p] "
               "[-s checkpoint_out] [-r checkpoint_in] [-n] [-a] [-g gpu]\n";
  return 2;
}

bool parse_args(int argc, char **argv, Args &a) {
  for (int i = 1; i < argc; ++i) {
    const std::string f = argv[i];
    auto need = [&](const char *what) -> const char * {
      if (i + 1 >= argc) {
        std::cerr << "mint: " << f << " needs " << what << "\n";
        return nullptr;
      }
      return argv[++i];
    };
    if (f == "-i") {
      a.info = true;
    } else if (f == "-p") {
      a.print_positions = true;
    } else if (f == "-n") {
      a.dry = true;
    } else if (f == "-a") {
      a.all_sweeps = true;
    } else if (f == "-c" || f == "-t" || f == "-o" || f == "-d" || f == "-s" || f == "-r" || f == "-g") {
      const char *v = need(f == "-t" || f == "-d" || f == "-g" ? "a number" : "a file");
      if (!v) return false;
      if (f == "-c") a.config = v;
      if (f == "-s") a.save = v;
      if (f == "-r") a.resume = v;
      if (f == "-o") a.output = v;
      if (f == "-t") a.steps = std::atoll(v);
      if (f == "-d") a.device = std::atoi(v);
      if (f == "-g") a.gpu = std::atoi(v);
    } else {
      std::cerr << "mint: unknown flag " << f << "\n";
      return false;
    }
  }
  return !a.config.empty() && a.steps >= 0 && a.device >= 0;
}

void check(int rc, void *ctx, const char *what) {
  if (rc != 0) {
    std::cerr << "mint: " << what << " failed (" << rc << "): " << mint_last_error(ctx) << "\n";
    std::exit(1);
  }
}

}  // namespace

int main(int argc, char **argv) {
  Args args;
  if (!parse_args(argc, argv, args)) return usage();

  std::ifstream in(args.config);
  if (!in) {
    std::cerr << "mint: cannot open " << args.config << "\n";
    return 1;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  jsonmin::Value cfgj;
  std::vector<Body> bodies, sweeps;
  try {
    cfgj = jsonmin::parse(ss.str());
    const jsonmin::Value &bl = cfgj.at("bodies");
    for (size_t i = 0; i < bl.arr.size(); ++i) bodies.push_back(body_from(bl.arr[i], "body-" + std::to_string(i)));
    if (cfgj.has("sweeps"))
      for (const auto &s : cfgj.at("sweeps").arr) sweeps.push_back(body_from(s, "sweep"));
  } catch (const std::exception &e) {
    std::cerr << "mint: " << args.config << ": " << e.what() << "\n";
    return 1;
  }
  if (bodies.empty()) {
    std::cerr << "mint: no bodies in " << args.config << "\n";
    return 1;
  }
  if (!sweeps.empty() && !args.all_sweeps) {
    const size_t first = (size_t)SWEEPS_PER_DEVICE * (size_t)args.device;
    if (first >= sweeps.size()) {
      std::cerr << "mint: -d " << args.device << " selects sweeps " << first << ".." << first + SWEEPS_PER_DEVICE - 1
                << " but " << args.config << " has " << sweeps.size() << " (use -a for all of them)\n";
      return 1;
    }
    const size_t last = std::min(sweeps.size(), first + SWEEPS_PER_DEVICE);
    sweeps = std::vector<Body>(sweeps.begin() + (long)first, sweeps.begin() + (long)last);
  }

  mint_config cfg;
  mint_config_default(&cfg);
  cfg.dt = cfgj.number_or("dt", cfg.dt);
  cfg.n_hill = cfgj.number_or("n_hill", cfg.n_hill);
  cfg.bs_tol = cfgj.number_or("bs_tol", cfg.bs_tol);
  cfg.softening2 = cfgj.number_or("softening2", cfg.softening2);
  cfg.central_mass = cfgj.number_or("central_mass", cfg.central_mass);
  const std::string form = cfgj.string_or("form", "batched");
  cfg.form = form == "plain" ? MINT_FORM_PLAIN : form == "flat" ? MINT_FORM_FLAT : MINT_FORM_BATCHED;
  cfg.device = args.gpu >= 0 ? args.gpu : args.device;
  const long long ckpt = (long long)cfgj.number_or("checkpoint_every", (double)args.steps);

  // systems: one per sweep (sweep replaces body 0), or the bodies as given
  const int S = sweeps.empty() ? 1 : (int)sweeps.size();
  const int NB = (int)bodies.size();
  // massive bodies (mass > 1e-12, src/mercurius.cu:9) first, original order kept
  std::vector<int> order;
  for (int pass = 0; pass < 2; ++pass)
    for (int i = 0; i < NB; ++i) {
      const double m = (i == 0 && !sweeps.empty()) ? sweeps[0].mass : bodies[i].mass;
      if ((m > 1e-12) == (pass == 0)) order.push_back(i);
    }
  int NM = 0;
  for (int i = 0; i < NB; ++i) {
    const double m = (i == 0 && !sweeps.empty()) ? sweeps[0].mass : bodies[i].mass;
    NM += m > 1e-12;
  }
  const int NT = NB - NM;
  if (NM < 1 || NM > 2048 || (NM > 8 && NT > 0)) {
    std::cerr << "mint: " << NM << " massive and " << NT << " massless bodies; this build integrates 1 to 8 "
                 "massive bodies per system with test particles, up to 2048 without\n";
    return 1;
  }
  for (const Body &s : sweeps)
    if (!(s.mass > 1e-12)) {
      std::cerr << "mint: sweep bodies must be massive\n";
      return 1;
    }

  const double mu = cfg.central_mass;
  std::vector<int64_t> off(S + 1);
  std::vector<double> m((size_t)S * NM), mx((size_t)S * NM * 3), mv((size_t)S * NM * 3);
  std::vector<double> t[6];
  for (auto &a : t) a.resize((size_t)S * NT);
  for (int s = 0; s < S; ++s) {
    off[s] = (int64_t)s * NT;
    for (int k = 0; k < NB; ++k) {
      const int i = order[k];
      const Body &b = (i == 0 && !sweeps.empty()) ? sweeps[s] : bodies[i];
      double x[3], v[3];
      check(mint_elements_to_cartesian(1, mu, &b.inc, &b.node, &b.argp, &b.manom, &b.e, &b.a, x, v), nullptr,
            "elements_to_cartesian");
      if (k < NM) {
        m[(size_t)s * NM + k] = b.mass;
        for (int d = 0; d < 3; ++d) {
          mx[((size_t)s * NM + k) * 3 + d] = x[d];
          mv[((size_t)s * NM + k) * 3 + d] = v[d];
        }
      } else {
        const size_t j = (size_t)s * NT + (k - NM);
        for (int d = 0; d < 3; ++d) {
          t[d][j] = x[d];
          t[3 + d][j] = v[d];
        }
      }
    }
  }
  off[S] = (int64_t)S * NT;

  if (args.dry) {
    double msum = 0.0;
    for (double v : m) msum += v;
    std::cout << "Config " << args.config << ": " << S << " system(s), " << NM << " massive and " << NT
              << " massless bodies each, dt " << cfg.dt << " yr, n_hill " << cfg.n_hill << ", bs_tol "
              << cfg.bs_tol << ", form " << form << ", total massive mass " << msum << "\n";
    if (args.info)
      for (int k = 0; k < NB; ++k) {
        const Body &b = bodies[order[k]];
        std::cout << "Body " << order[k] << " " << b.name << ": mass " << b.mass << " a " << b.a << " e " << b.e
                  << " i " << b.inc << "\n";
      }
    return 0;
  }
  void *ctx = nullptr;
  if (mint_create(&cfg, &ctx) != 0) {
    std::cerr << "mint: " << mint_last_error(nullptr) << "\n";
    return 1;
  }
  mint_state st{S, NM, off.data(), m.data(), mx.data(), mv.data(), t[0].data(), t[1].data(), t[2].data(),
                t[3].data(), t[4].data(), t[5].data(), nullptr};
  check(mint_upload(ctx, &st), ctx, "upload");
  // physical radii: per-body "radius" [AU], else from "density" [g/cm^3]
  const double density = cfgj.number_or("density", -1.0);
  bool any_radius = density > 0.0;
  for (const Body &b : bodies) any_radius = any_radius || b.radius >= 0.0;
  for (const Body &b : sweeps) any_radius = any_radius || b.radius >= 0.0;
  if (any_radius) {
    const double au_cm = 1.495978707e13, msun_g = 1.98847e33;
    const double rho = density > 0.0 ? density * au_cm * au_cm * au_cm / msun_g : -1.0;  // Msun / AU^3
    std::vector<double> rad((size_t)S * NM, 0.0);
    for (int s = 0; s < S; ++s)
      for (int k = 0; k < NM; ++k) {
        const int i = order[k];
        const Body &b = (i == 0 && !sweeps.empty()) ? sweeps[s] : bodies[i];
        rad[(size_t)s * NM + k] = b.radius >= 0.0 ? b.radius
                                  : rho > 0.0    ? std::cbrt(3.0 * b.mass / (4.0 * M_PI * rho))
                                                 : 0.0;
      }
    check(mint_set_radii(ctx, rad.data()), ctx, "radii");
  }
  int64_t step0 = 0;
  if (!args.resume.empty()) {
    check(mint_load_checkpoint(ctx, args.resume.c_str(), &step0), ctx, "resume");
    if (args.info) std::cout << "Resumed from " << args.resume << " at step " << step0 << "\n";
  }
  std::vector<double> E0(S), E1(S);
  check(mint_energy(ctx, E0.data()), ctx, "energy");
  if (args.info) {
    std::cout << "Simulation: " << S << " system(s), " << NM << " massive and " << NT << " massless bodies each, dt "
              << cfg.dt << " yr, " << args.steps << " steps, form " << form << "\n";
    for (int k = 0; k < NB; ++k) {
      const Body &b = bodies[order[k]];
      std::cout << "Body " << order[k] << " " << b.name << ": mass " << b.mass << " a " << b.a << " e " << b.e
                << " i " << b.inc << "\n";
    }
  }
  if (cfg.form == MINT_FORM_BATCHED && (long long)S * NT >= 100000) {
    int32_t best[3];
    check(mint_autotune(ctx, nullptr, best), ctx, "autotune");
    if (args.info) std::cout << "Launch shape: block " << best[0] << ", TPs/lane " << best[1] << ", window " << best[2] << "\n";
  }

  const std::string pos_file = args.output.empty() ? "output_" + std::to_string(args.device) : args.output;
  long long done = 0;
  std::vector<double> ox(mx.size()), ov(mv.size());
  std::vector<double> o[6];
  for (auto &a : o) a.resize(t[0].size());
  mint_state out{S, NM, off.data(), nullptr, ox.data(), ov.data(), o[0].data(), o[1].data(), o[2].data(),
                 o[3].data(), o[4].data(), o[5].data(), nullptr};
  while (done < args.steps) {
    const long long n = std::min(ckpt > 0 ? ckpt : args.steps, args.steps - done);
    check(mint_step(ctx, n), ctx, "step");
    done += n;
    check(mint_download(ctx, &out), ctx, "download");
    // positions (write_positions format, src/simutils.cuh:186-230)
    std::ofstream pf(pos_file, std::ios::app);
    if (!pf) {
      std::cerr << "mint: cannot open " << pos_file << "\n";
      return 1;
    }
    int nan = 0;
    pf << "#\n";
    std::vector<std::vector<double>> xyz((size_t)S * NB, std::vector<double>(6));
    for (int s = 0; s < S; ++s) {
      pf << "$\n";
      for (int i = 0; i < NB; ++i) {
        int k = 0;
        while (order[k] != i) ++k;
        double *q = xyz[(size_t)s * NB + i].data();
        for (int d = 0; d < 3; ++d) {
          q[d] = k < NM ? ox[((size_t)s * NM + k) * 3 + d] : o[d][(size_t)s * NT + (k - NM)];
          q[3 + d] = k < NM ? ov[((size_t)s * NM + k) * 3 + d] : o[3 + d][(size_t)s * NT + (k - NM)];
        }
        if (std::isnan(q[0]) || std::isnan(q[1]) || std::isnan(q[2])) ++nan;
        pf << q[0] << " " << q[1] << " " << q[2] << "\n";
        if (args.print_positions) std::cout << s << " " << i << " " << q[0] << " " << q[1] << " " << q[2] << "\n";
      }
      pf << "\n";
    }
    pf << "\n";
    std::cout << "NaN counter after step " << step0 + done << ": " << nan << std::endl;
    // elements (src/mercurius.cu:210-248)
    std::ofstream ef("elements_" + std::to_string(args.device) + ".txt");
    ef << "Timestep: " << step0 + done << "\n";
    for (int s = 0; s < S; ++s) {
      ef << "Experiment: " << s << "\n";
      for (int i = 0; i < NB; ++i) {
        const double *q = xyz[(size_t)s * NB + i].data();
        double el[6];
        mint_cartesian_to_elements(1, mu, q, q + 3, &el[0], &el[1], &el[2], &el[3], &el[4], &el[5]);
        ef << el[0] << " " << el[1] << " " << el[2] << " " << el[3] << " " << el[4] << " " << el[5] << "\n";
      }
      ef << "\n";
    }
  }
  if (!args.save.empty()) check(mint_save_checkpoint(ctx, args.save.c_str(), step0 + done), ctx, "checkpoint");
  check(mint_energy(ctx, E1.data()), ctx, "energy");
  std::vector<uint32_t> status(S);
  check(mint_status(ctx, status.data()), ctx, "status");
  double worst = 0.0;
  uint32_t any = 0;
  for (int s = 0; s < S; ++s) {
    worst = std::max(worst, std::fabs((E1[s] - E0[s]) / E0[s]));
    any |= status[s];
  }
  std::cout << "Energy drift max |dE/E| over systems: " << worst << "; status bits: " << any << std::endl;
  mint_destroy(ctx);
  return 0;
}
