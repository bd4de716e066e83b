// spm -- the command-line program: a run from a UDF input, as the
// reference's main() (sp_3d_ns.cxx:392-640) with file_get's flags
// (input.cxx:1819-1877), in C++ over the library's C ABI only
// (include/spm.h, include/spm_udf.h).
//
//   spm -I input.udf -O output.udf -R restart.udf [-D define.udf]
//       [--state FILE] [--resume-state FILE] [--fp32] [--schedule literal] [--stress FILE] [--quiet]
//
// Values may be attached (-Iinput.udf) as the reference takes them.  The
// time loop is the reference's: steps ts = 0..MSTEP (MSTEP = GTS * Num_snap),
// a snapshot record before the step whenever ts % GTS == 0, the restart
// after step MSTEP.  An input whose resume.Calculation is CONTINUE resumes:
// one step from the fresh initial state, then the saved state replaces it
// (sp_3d_ns.cxx:546-551, 619-628) and the loop goes on from the saved ts.
// The library counts ts after each step, the reference's jikan.ts is the
// index of the step running: the restart holds ts = (steps done) - 1 and a
// restored context is given ts + 1.  --state writes the exact binary
// checkpoint (spm_save_state) after the last step, --resume-state starts
// from one (no counterpart in the reference).  Under Lees-Edwards shear
// --stress FILE writes Mean_shear_stress's line after every step, as the
// reference prints it to stderr (sp_3d_ns.h:209-358, sp_3d_ns.cxx:598-601).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spm.h"
#include "../../include/spm_udf.h"

namespace {

struct Args {
  std::string input, output, restart, define, state, resume_state, stress;
  bool fp32 = false, literal = false, quiet = false;
};

[[noreturn]] void die(const std::string &msg) {
  std::fprintf(stderr, "spm: %s\n", msg.c_str());
  std::exit(1);
}

void check(int rc, const char *what, spm_ctx *ctx = nullptr) {
  if (rc == SPM_OK) return;
  const char *detail = ctx ? spm_last_error(ctx) : spm_last_error(nullptr);
  die(std::string(what) + ": " + spm_strerror(rc) + (detail && *detail ? std::string(": ") + detail : ""));
}

void check_udf(int rc, const char *what) {
  if (rc != SPM_OK) die(std::string(what) + ": " + spm_strerror(rc) + ": " + spm_udf_last_error());
}

Args parse(int argc, char **argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    auto value = [&](std::string &dst) {
      if (s.size() > 2) {
        dst = s.substr(2);
      } else {
        if (i + 1 >= argc) die("flag " + s + " needs a value");
        dst = argv[++i];
      }
    };
    auto long_value = [&](std::string &dst) {
      if (i + 1 >= argc) die("flag " + s + " needs a value");
      dst = argv[++i];
    };
    if (s == "--state") long_value(a.state);
    else if (s == "--resume-state") long_value(a.resume_state);
    else if (s == "--stress") long_value(a.stress);
    else if (s == "--fp32") a.fp32 = true;
    else if (s == "--quiet") a.quiet = true;
    else if (s == "--schedule") {
      std::string v;
      long_value(v);
      if (v != "literal" && v != "fused") die("--schedule takes fused or literal");
      a.literal = v == "literal";
    } else if (s.size() >= 2 && s[0] == '-' && s[1] != '-') {
      std::string dummy;
      switch (s[1]) {
        case 'I': value(a.input); break;
        case 'O': value(a.output); break;
        case 'R': value(a.restart); break;
        case 'D': value(a.define); break;
        case 'S': case 'M': value(dummy); break;  // accepted, unused (as file_get's summary/control)
        default: die("unknown flag " + s);
      }
    } else {
      die("unknown argument " + s);
    }
  }
  if (a.input.empty() || a.output.empty() || a.restart.empty())
    die("usage: spm -I input.udf -O output.udf -R restart.udf [-D define.udf] [--state FILE] "
        "[--resume-state FILE] [--fp32] [--schedule fused|literal] [--stress FILE] [--quiet]");
  return a;
}

bool string_is(const spm_udf *u, const char *loc, const char *want) {
  char buf[256] = {};
  if (spm_udf_get_string(u, loc, buf, sizeof buf) != SPM_OK) return false;
  return std::strcmp(buf, want) == 0;
}

size_t field_len(const spm_params &p) { return (size_t)p.nx * p.ny * 2 * (p.nz / 2 + 1); }

}  // namespace

int main(int argc, char **argv) {
  const Args a = parse(argc, argv);
  auto log = [&](const char *fmt, auto... v) {
    if (!a.quiet) {
      std::fprintf(stderr, fmt, v...);
      std::fputc('\n', stderr);
    }
  };

  spm_udf *u = nullptr;
  check_udf(spm_udf_open(a.input.c_str(), &u), "spm_udf_open");
  spm_params prm{};
  spm_electrolyte el{};
  spm_udf_run run{};
  check_udf(spm_udf_params(u, &prm, &el, &run), "spm_udf_params");
  const bool electro = run.equation == SPM_UDF_ELECTROLYTE;
  const bool sheared = run.equation == SPM_UDF_LEES_EDWARDS;
  if (electro && run.poisson_boltzmann && !(el.n_spec == 2 && prm.n_particles > 1))
    die("INIT_profile Poisson_Boltzmann is run for salt with several particles only");
  if (a.fp32) prm.precision = SPM_FP32;
  const int gts = run.gts, mstep = run.gts * run.num_snap;
  const bool resumed = string_is(u, "resume.Calculation", "CONTINUE");
  const bool write_udf = string_is(u, "output.UDF", "ON");
  const int n_ion = electro ? el.n_spec : 0;
  log("#mesh = %d * %d * %d, particles %d, DT %.17g, MSTEP %d, GTS %d", prm.nx, prm.ny, prm.nz,
      prm.n_particles, prm.dt, mstep, gts);

  spm_ctx *ctx = nullptr;
  check(spm_create(&prm, &ctx), "spm_create");
  check(spm_set_schedule(ctx, a.literal ? SPM_SCHEDULE_LITERAL : SPM_SCHEDULE_FUSED), "spm_set_schedule", ctx);

  // Init_Particle and the MT19937 state it leaves (the thermostat goes on
  // from it); a resumed run of another distribution takes only the species
  // (the restart replaces the rest)
  const int np = prm.n_particles;
  std::vector<spm_particle> p((size_t)(np > 0 ? np : 1));
  std::memset(p.data(), 0, sizeof(spm_particle) * p.size());
  const int d = run.distribution;
  const bool placed = d == SPM_UDF_USER_SPECIFY || d == SPM_UDF_UNIFORM_RANDOM || d == SPM_UDF_BCC ||
                      d == SPM_UDF_FCC || d == SPM_UDF_RANDOM_WALK;
  if (placed) {
    std::vector<uint32_t> mt(SPM_MT_N);
    int mti = 0;
    check_udf(spm_udf_init_particles_rng(u, p.data(), np, mt.data(), &mti), "spm_udf_init_particles");
    check(spm_put_particles(ctx, p.data(), np), "spm_put_particles", ctx);
    check(spm_put_rng(ctx, mt.data(), mti), "spm_put_rng", ctx);
  } else {
    if (!resumed && a.resume_state.empty()) die("this INIT_distribution is not run");
    int k = 0;
    const double L[3] = {prm.nx * prm.dx, prm.ny * prm.dx, prm.nz * prm.dx};
    for (int s = 0; s < prm.n_species; ++s) {
      char loc[128];
      std::snprintf(loc, sizeof loc, "object_type.spherical_particle.Particle_spec[%d].Particle_number", s);
      int cnt = 0;
      check_udf(spm_udf_get_int(u, loc, &cnt), "Particle_number");
      for (int i = 0; i < cnt && k < np; ++i, ++k) p[k].spec = s;
    }
    for (int n = 0; n < np; ++n) {  // placeholders inside the box
      p[n].q.s = p[n].q_old.s = 1.0;
      for (int c = 0; c < 3; ++c) p[n].x[c] = p[n].x_nopbc[c] = (n + 0.5) / np * L[c];
    }
    check(spm_put_particles(ctx, p.data(), np), "spm_put_particles", ctx);
  }
  check(spm_init_fluid(ctx), "spm_init_fluid", ctx);
  if (electro) {
    check(spm_set_electrolyte(ctx, &el), "spm_set_electrolyte", ctx);
    if (run.poisson_boltzmann) {
      int iters = 0;
      check(spm_init_rho_ion_pb(ctx, &iters), "spm_init_rho_ion_pb", ctx);
      log("# Poisson-Boltzmann start: %d iterations", iters);
    } else {
      check(spm_init_rho_ion(ctx), "spm_init_rho_ion", ctx);
    }
  }
  if (sheared) check(spm_set_lees_edwards(ctx, run.shear_rate), "spm_set_lees_edwards", ctx);
  if (write_udf) check_udf(spm_udf_create_output(a.input.c_str(), a.output.c_str()), "spm_udf_create_output");

  const size_t fl = field_len(prm);
  std::vector<double> z0(fl), z1(fl), conc(fl * (n_ion > 0 ? n_ion : 1));
  int ts = 0;
  if (!a.resume_state.empty()) {
    check(spm_load_state(ctx, a.resume_state.c_str()), "spm_load_state", ctx);
    double t = 0.0;
    check(spm_get_time(ctx, &ts, &t), "spm_get_time", ctx);
  } else if (resumed) {
    check(spm_step(ctx, 1), "spm_step", ctx);  // the first loop iteration runs before the restore
    check(spm_get_particles(ctx, p.data(), np), "spm_get_particles", ctx);
    spm_restart r{};
    r.nx = prm.nx;
    r.ny = prm.ny;
    r.nz = prm.nz;
    r.n_particles = np;
    r.n_ion = n_ion;
    r.zeta0 = z0.data();
    r.zeta1 = z1.data();
    r.concentration = conc.data();
    r.particles = p.data();
    check_udf(spm_udf_read_restart(u, &r), "spm_udf_read_restart");
    if (sheared) {  // only degree_oblique is restored (resume.cxx:335-339)
      double deg = 0.0, rate = 0.0;
      check(spm_get_oblique(ctx, &deg, &rate), "spm_get_oblique", ctx);
      check(spm_put_oblique(ctx, r.degree_oblique, rate), "spm_put_oblique", ctx);
    }
    check(spm_put_field(ctx, SPM_FIELD_ZETA0, z0.data()), "spm_put_field", ctx);
    check(spm_put_field(ctx, SPM_FIELD_ZETA1, z1.data()), "spm_put_field", ctx);
    check(spm_put_uk_dc(ctx, r.uk_dc), "spm_put_uk_dc", ctx);
    check(spm_put_particles(ctx, p.data(), np), "spm_put_particles", ctx);
    for (int n = 0; n < n_ion; ++n)
      check(spm_put_concentration(ctx, n, conc.data() + fl * n), "spm_put_concentration", ctx);
    check(spm_put_time(ctx, r.ts + 1, r.time), "spm_put_time", ctx);
    log("############################ Parameters are restored.");
    ts = r.ts + 1;
  }
  FILE *fstress = nullptr;
  if (!a.stress.empty()) {
    if (!sheared) die("--stress needs Lees-Edwards input");
    fstress = std::fopen(a.stress.c_str(), "w");
    if (!fstress) die("cannot write " + a.stress);
    std::fprintf(fstress, "#1:time 2:shear_rate 3:shear_strain_temporal 4:lj_dev_stress_temporal "
                          "5:shear_stress_temporal_old 6:shear_stress_temporal_new 7:viscosity \n");
  }
  while (ts <= mstep) {
    if (ts % gts == 0) {
      if (write_udf) {
        check(spm_get_particles(ctx, p.data(), np), "spm_get_particles", ctx);
        check_udf(spm_udf_append_record(a.output.c_str(), ts, p.data(), np), "spm_udf_append_record");
      }
      log("# Step: %9d/%9d", ts, mstep);
    }
    int next = (ts / gts + 1) * gts < mstep + 1 ? (ts / gts + 1) * gts : mstep + 1;
    if (fstress) next = ts + 1;  // a line after every step
    check(spm_step(ctx, next - ts), "spm_step", ctx);
    ts = next;
    if (fstress) {
      spm_le_stress st{};
      check(spm_le_shear_stress(ctx, &st), "spm_le_shear_stress", ctx);
      std::fprintf(fstress, "%g %g %g %g %g %g %g\n", st.time, st.shear_rate, st.strain, st.lj_dev_stress,
                   st.hydro_stress, st.hydro_stress_new, st.viscosity);
    }
  }
  if (fstress) std::fclose(fstress);

  int ts_done = 0;
  double time = 0.0;
  check(spm_get_time(ctx, &ts_done, &time), "spm_get_time", ctx);
  check(spm_get_field(ctx, SPM_FIELD_ZETA0, z0.data()), "spm_get_field", ctx);
  check(spm_get_field(ctx, SPM_FIELD_ZETA1, z1.data()), "spm_get_field", ctx);
  check(spm_get_particles(ctx, p.data(), np), "spm_get_particles", ctx);
  for (int n = 0; n < n_ion; ++n)
    check(spm_get_concentration(ctx, n, conc.data() + fl * n), "spm_get_concentration", ctx);
  spm_restart r{};
  r.nx = prm.nx;
  r.ny = prm.ny;
  r.nz = prm.nz;
  r.n_particles = np;
  r.n_ion = n_ion;
  r.ts = ts_done - 1;
  r.time = time;
  check(spm_get_uk_dc(ctx, r.uk_dc), "spm_get_uk_dc", ctx);
  if (sheared) {
    double rate = 0.0;
    check(spm_get_oblique(ctx, &r.degree_oblique, &rate), "spm_get_oblique", ctx);
  }
  r.zeta0 = z0.data();
  r.zeta1 = z1.data();
  r.concentration = n_ion ? conc.data() : nullptr;
  r.particles = p.data();
  check_udf(spm_udf_write_restart(a.input.c_str(), a.restart.c_str(), &r), "spm_udf_write_restart");
  if (!a.state.empty()) check(spm_save_state(ctx, a.state.c_str()), "spm_save_state", ctx);
  spm_destroy(ctx);
  spm_udf_close(u);
  return 0;
}
