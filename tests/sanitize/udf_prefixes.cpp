// Sanitizer harness for the UDF reader (kapsel_old_amd/csrc/udf.cpp), built
// on the CPU with -fsanitize=address,undefined by tests/test_udf_sanitized.py.
//
//   udf_prefixes STEP DIR FILE...
//
// For every FILE (a .udf whose \include targets sit in DIR, a scratch
// directory holding copies of the file's companions), parses the whole
// file and then every prefix cut at a multiple of STEP bytes (and at each of
// the last 64 bytes), written over DIR/basename; on a file that opens it
// reads the parameters, the particles and the records as the library's users
// do.  A prefix may fail to parse; it must not read out of bounds, leak or
// hang.  Prints one line per file.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/spm_udf.h"

static std::string slurp(const char *path) {
  std::ifstream f(path, std::ios::binary);
  std::ostringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

static void spit(const std::string &path, const std::string &text, size_t n) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  f.write(text.data(), (std::streamsize)n);
}

static void exercise(spm_udf *u) {
  spm_params prm{};
  spm_electrolyte el{};
  spm_udf_run run{};
  if (spm_udf_params(u, &prm, &el, &run) == SPM_OK && prm.n_particles > 0 && prm.n_particles <= 5000) {
    std::vector<spm_particle> p((size_t)prm.n_particles);
    (void)spm_udf_init_particles(u, p.data(), prm.n_particles);
  }
  int nrec = 0;
  if (spm_udf_num_records(u, &nrec) == SPM_OK)
    for (int r = 0; r < nrec && r < 4; ++r) (void)spm_udf_jump(u, r);
  char buf[64];
  (void)spm_udf_get_string(u, "resume.Calculation", buf, sizeof buf);
  double xs[16];
  int n = 0;
  (void)spm_udf_get_doubles(u, "Particles[].R.x", xs, 16, &n);
}

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: udf_prefixes STEP DIR FILE...\n");
    return 2;
  }
  const size_t step = (size_t)std::atol(argv[1]);
  const std::string dir = argv[2];
  for (int i = 3; i < argc; ++i) {
    const std::string text = slurp(argv[i]);
    std::string base = argv[i];
    base = base.substr(base.find_last_of('/') + 1);
    const std::string path = dir + "/" + base;
    int ok = 0, bad = 0;
    std::vector<size_t> cuts;
    for (size_t c = 0; c < text.size(); c += step) cuts.push_back(c);
    for (size_t c = text.size() > 64 ? text.size() - 64 : 0; c <= text.size(); ++c) cuts.push_back(c);
    for (size_t c : cuts) {
      spit(path, text, c);
      spm_udf *u = nullptr;
      if (spm_udf_open(path.c_str(), &u) == SPM_OK) {
        ++ok;
        if (c == text.size()) exercise(u);
        spm_udf_close(u);
      } else {
        ++bad;
      }
    }
    spit(path, text, text.size());  // restore the companion
    std::printf("%s: %zu bytes, %d prefixes parsed, %d rejected\n", base.c_str(), text.size(), ok, bad);
  }
  return 0;
}
