// checkpoint.cpp -- the context's state to a file and back, exactly.
//
// The reference resumes from the restart UDF (Save_Restart_udf /
// Force_restore_parameters, resume.cxx:10-364), which holds zeta, uk_dc, the
// particles and the time but not the Mersenne Twister state: a resumed run
// with kBT > 0 draws different noise (the generator is reseeded at start,
// sp_3d_ns.cxx:477).  This file holds everything the next step reads -- the
// two vorticity spectra in the context's own precision, uk_dc, every
// particle record, ts and time, the MT19937 words and index, and for an
// electrolyte the ion spectra and their conserved totals, and under
// Lees-Edwards shear degree_oblique and Shear_rate_eff (the reference's
// restart keeps only degree_oblique, resume.cxx:335-339, and takes
// Shear_rate_eff from the first step) -- so that a context loaded from it
// continues bit for bit like the one that saved it.
//
// Layout (native byte order): the header below, then zeta0, zeta1
// (NX*NY*2(NZ/2+1) elements each in the reference layout, 4 or 8 bytes), the particle records
// (n_particles * sizeof(spm_particle)), the ion spectra (n_ion * nfield
// elements), and an FNV-1a 64-bit hash of every byte before it.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "spm_internal.h"

namespace {

constexpr char kMagic[8] = {'S', 'P', 'M', 'S', 'T', 'A', 'T', 'E'};
constexpr uint32_t kVersion = 2;

struct Header {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  int32_t nx, ny, nz, precision;
  int32_t n_particles, n_ion, ts, mti;
  double time, dt;
  double uk_dc[3];
  double total_solute[2];
  int32_t le, le_pad;  // Lees-Edwards shear: 1 and the three values below
  double le_rate, le_deg, le_sre;
  uint32_t mt[SPM_MT_N];
};

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 1099511628211ull;
    }
  }
};

int fail(spm_ctx *c, int rc, const std::string &msg) {
  c->err = msg;
  return rc;
}

size_t elem_bytes(const spm_ctx *c) { return c->prm.precision == SPM_FP32 ? 4 : 8; }

// device buffers saved, in file order
std::vector<void *> buffers(spm_ctx *c) {
  std::vector<void *> b = {c->zeta[0], c->zeta[1]};
  if (c->electro)
    for (int n = 0; n < c->el.n_spec; ++n) b.push_back(c->conc[n]);
  return b;
}

}  // namespace

extern "C" int spm_save_state(spm_ctx *c, const char *path) {
  if (!c || !path) return SPM_ERR_INVALID;
  c->err.clear();
  SPM_HIP(c, hipSetDevice(c->device));
  Header h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.header_bytes = sizeof(Header);
  h.nx = c->prm.nx;
  h.ny = c->prm.ny;
  h.nz = c->prm.nz;
  h.precision = c->prm.precision;
  h.n_particles = c->np;
  h.n_ion = c->electro ? c->el.n_spec : 0;
  h.ts = c->ts;
  h.mti = c->mti;
  h.time = c->time;
  h.dt = c->prm.dt;
  for (int n = 0; n < 2; ++n) h.total_solute[n] = c->total_solute[n];
  std::memcpy(h.mt, c->mt, sizeof(h.mt));
  SPM_HIP(c, hipMemcpyAsync(h.uk_dc, c->uk_dc, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
  if (c->le) {
    h.le = 1;
    h.le_rate = c->le_rate;
    h.le_deg = c->le_deg;
    SPM_HIP(c, hipMemcpyAsync(&h.le_sre, c->le_red + spm::le_red_slots() - 2, sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
  }
  // fields in the reference layout (independent of the device row pitch)
  const size_t fb = elem_bytes(c) * spm::ref_count(c->grid);
  std::vector<std::vector<char>> data;
  for (void *b : buffers(c)) {
    data.emplace_back(fb);
    if (int rc = spm::rows_to_host(c, data.back().data(), b, elem_bytes(c))) return rc;
  }
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<spm_particle> p(c->np > 0 ? c->np : 1);
  if (int rc = spm_get_particles(c, p.data(), c->np)) return rc;

  const std::string tmp = std::string(path) + ".tmp";
  std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
  if (!f) return fail(c, SPM_ERR_INVALID, "cannot write " + tmp);
  Fnv hash;
  auto put = [&](const void *src, size_t n) {
    f.write(static_cast<const char *>(src), (std::streamsize)n);
    hash.add(src, n);
  };
  put(&h, sizeof(h));
  put(data[0].data(), fb);
  put(data[1].data(), fb);
  put(p.data(), sizeof(spm_particle) * (size_t)c->np);
  for (size_t k = 2; k < data.size(); ++k) put(data[k].data(), fb);
  f.write(reinterpret_cast<const char *>(&hash.h), sizeof(hash.h));
  f.close();
  if (!f) return fail(c, SPM_ERR_INVALID, "write failed on " + tmp);
  if (std::rename(tmp.c_str(), path) != 0) return fail(c, SPM_ERR_INVALID, std::string("cannot rename to ") + path);
  return SPM_OK;
}

extern "C" int spm_load_state(spm_ctx *c, const char *path) {
  if (!c || !path) return SPM_ERR_INVALID;
  c->err.clear();
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(c, SPM_ERR_INVALID, std::string("cannot read ") + path);
  Fnv hash;
  auto get = [&](void *dst, size_t n) -> bool {
    f.read(static_cast<char *>(dst), (std::streamsize)n);
    if ((size_t)f.gcount() != n) return false;
    hash.add(dst, n);
    return true;
  };
  Header h;
  if (!get(&h, sizeof(h)) || std::memcmp(h.magic, kMagic, 8) != 0)
    return fail(c, SPM_ERR_INVALID, std::string(path) + " is not a state file");
  if (h.version != kVersion || h.header_bytes != sizeof(Header))
    return fail(c, SPM_ERR_INVALID, "state file version differs");
  const int n_ion = c->electro ? c->el.n_spec : 0;
  if (h.nx != c->prm.nx || h.ny != c->prm.ny || h.nz != c->prm.nz || h.precision != c->prm.precision ||
      h.n_particles != c->np || h.n_ion != n_ion || h.dt != c->prm.dt || h.le != (c->le ? 1 : 0) ||
      (c->le && h.le_rate != c->le_rate))
    return fail(c, SPM_ERR_INVALID, "state file was saved by a context of another grid, precision, particle "
                                    "count, ion count, time step or shear");
  if (h.ts < 0 || h.mti < 0 || h.mti > SPM_MT_N + 1) return fail(c, SPM_ERR_INVALID, "corrupt state header");
  const size_t fb = elem_bytes(c) * spm::ref_count(c->grid);
  std::vector<std::vector<char>> data(2 + n_ion, std::vector<char>(fb));
  std::vector<spm_particle> p(c->np > 0 ? c->np : 1);
  bool ok = get(data[0].data(), fb) && get(data[1].data(), fb) && get(p.data(), sizeof(spm_particle) * (size_t)c->np);
  for (int k = 0; ok && k < n_ion; ++k) ok = get(data[2 + k].data(), fb);
  uint64_t stored = 0;
  if (!ok || !f.read(reinterpret_cast<char *>(&stored), sizeof(stored)) || stored != hash.h)
    return fail(c, SPM_ERR_INVALID, std::string(path) + " is truncated or corrupt");
  // everything is checked before the context changes
  if (int rc = spm_put_particles(c, p.data(), c->np)) return rc;
  SPM_HIP(c, hipSetDevice(c->device));
  const std::vector<void *> dst = buffers(c);
  for (size_t k = 0; k < dst.size(); ++k)
    if (int rc = spm::rows_to_device(c, dst[k], data[k].data(), elem_bytes(c))) return rc;
  SPM_HIP(c, hipMemcpyAsync(c->uk_dc, h.uk_dc, sizeof(double) * 3, hipMemcpyHostToDevice, c->stream));
  if (c->le)
    SPM_HIP(c, hipMemcpyAsync(c->le_red + spm::le_red_slots() - 2, &h.le_sre, sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  if (c->le) c->le_deg = h.le_deg;
  c->ts = h.ts;
  c->time = h.time;
  std::memcpy(c->mt, h.mt, sizeof(c->mt));
  c->mti = h.mti;
  for (int n = 0; n < 2; ++n) c->total_solute[n] = h.total_solute[n];
  return SPM_OK;
}
