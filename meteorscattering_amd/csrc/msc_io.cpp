// Parameters and run-directory files (the reference's process contract).
//
// inputs.dat reader: meteor-fork.cpp:84-115 (writer setup/meteorrun.m:45-69)
// B0.dat:            meteor-fork.cpp:221-226 (writer meteorrun.m:74-78)
// wp.dat / nu.dat:   meteor-fork.cpp:340-347 (writer meteorrun.m:108-114)
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "msc_common.h"

namespace msc {

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

namespace {

struct Reader {
  std::vector<unsigned char> b;
  size_t off = 0;
  bool get(void* dst, size_t n) {
    if (off + n > b.size()) return false;
    std::memcpy(dst, b.data() + off, n);
    off += n;
    return true;
  }
};

bool slurp(const std::string& path, std::vector<unsigned char>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(n > 0 ? (size_t)n : 0);
  size_t got = n > 0 ? std::fread(out.data(), 1, (size_t)n, f) : 0;
  std::fclose(f);
  return got == out.size();
}

bool spit(const std::string& path, const void* data, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  size_t w = n ? std::fwrite(data, 1, n, f) : 0;
  return std::fclose(f) == 0 && w == n;
}

const char kExtMagic[4] = {'M', 'S', 'C', 'X'};
const int32_t kExtVersion = 2;  // v1 lacks j_collocated

}  // namespace
}  // namespace msc

using namespace msc;

extern "C" {

const char* msc_last_error(void) { return g_last_error.c_str(); }

const char* msc_version(void) { return "meteorscattering_amd 0.1 (gfx950)"; }

void msc_free(void* p) { std::free(p); }

void msc_params_default(msc_params* p) {
  // setup/meteordefaults.m:1-92
  std::memset(p, 0, sizeof *p);
  p->dopml = 1;
  p->savefields = 1;
  p->f0 = 150e6;
  p->fsig = 200;
  p->sourcetype = 2;
  p->sourcepol = 3;
  p->gridfactor = 200;
  p->nx = p->ny = p->nz = 201;
  p->metsize[0] = p->metsize[1] = 0.5;
  p->wpmax = 1;
  p->numax = 1e7;
  p->numfiles = 100;
  p->nprobes = 1;
  p->probex[0] = p->probey[0] = p->probez[0] = 101;
  p->nntf = 15;
  for (int m = 0; m < 15; ++m) p->fntf[m] = (0.9 + 0.2 * m / 14.0) * p->f0;
  p->platesize = 1;
  p->B0[2] = 5e-5;
  double lambda = kC / p->f0, dx = lambda / p->gridfactor;
  double vp = 1 / std::sqrt(4 * kPi * 1e-7 * 8.845e-12);
  double stab = 1 / std::sqrt(1 + std::pow(p->wpmax * kPi / p->gridfactor / 2, 2));
  p->dt = dx / vp / std::sqrt(3.0) * stab * 0.8;  // meteorrun.m:33-34
  p->tsteps = (int)std::lround(p->nx * dx / vp / p->dt * 5);  // tfactor 5, :35
}

int msc_params_validate(const msc_params* p) {
  if (!p) return fail(MSC_EINVAL, "params is NULL");
  const int L = pml_len(p);
  if (L < 1 || L > 64) return fail(MSC_EINVAL, "pmllen %d out of range [1,64]", L);
  const int nmin = 2 * L + 22;  // CPML on both sides plus a non-empty TF/SF box and NTF planes
  if (p->nx < nmin || p->ny < nmin || p->nz < nmin)
    return fail(MSC_EINVAL, "grid %dx%dx%d too small: each side needs >= %d cells with pmllen %d", p->nx, p->ny,
                p->nz, nmin, L);
  if ((double)p->nx * p->ny * p->nz > 4.0e10) return fail(MSC_EINVAL, "grid too large");
  if (p->tsteps < 1) return fail(MSC_EINVAL, "tsteps must be >= 1 (got %d)", p->tsteps);
  if (p->numfiles < 1 || p->numfiles > p->tsteps)
    return fail(MSC_EINVAL, "numfiles must be in [1, tsteps] (got %d; the reference divides by tsteps/numfiles)",
                p->numfiles);
  if (p->nntf < 1 || p->nntf > MSC_MAX_NTF) return fail(MSC_EINVAL, "nntf %d out of range [1,%d]", p->nntf, MSC_MAX_NTF);
  if (p->nprobes < 0 || p->nprobes > MSC_MAX_PROBES)
    return fail(MSC_EINVAL, "nprobes %d out of range [0,%d]", p->nprobes, MSC_MAX_PROBES);
  if (p->j_collocated != 0 && p->j_collocated != 1)
    return fail(MSC_EINVAL, "j_collocated must be 0 or 1 (got %d)", p->j_collocated);
  for (int n = 0; n < p->nprobes; ++n)
    if (p->probex[n] < 1 || p->probex[n] >= p->nx || p->probey[n] < 1 || p->probey[n] >= p->ny || p->probez[n] < 1 ||
        p->probez[n] >= p->nz)
      return fail(MSC_EINVAL, "probe %d at (%d,%d,%d) outside [1,N) (the probe averages read index-1)", n,
                  p->probex[n], p->probey[n], p->probez[n]);
  if (!(p->dt > 0) || !(p->f0 > 0) || !(p->gridfactor > 0))
    return fail(MSC_EINVAL, "dt, f0 and gridfactor must be positive");
  if (p->doplate < 0 || p->doplate > 2) return fail(MSC_EINVAL, "doplate must be 0, 1 or 2");
  if (p->doplate == 1) {
    const double dx = (kC / p->f0) / p->gridfactor;
    const int ssx = (p->nx + 1) / 2, ssy = (p->ny + 1) / 2;
    const int lo = (int)(ssy - p->platesize / 2 / dx), up = (int)(ssy + p->platesize / 2 / dx);
    if (lo < 0 || up > p->ny || up > p->nz || ssx >= p->nx)
      return fail(MSC_EINVAL, "plate of size %g m does not fit the grid", p->platesize);
  }
  if (p->sourcepol < 1 || p->sourcepol > 5) {
    // the reference falls back to linear Ez for unknown values; so do we
  }
  return MSC_OK;
}

int msc_read_run_dir(const char* dir, msc_params* p, double** wp, double** nu) {
  if (!dir || !p || !wp || !nu) return fail(MSC_EINVAL, "NULL argument");
  *wp = *nu = nullptr;
  std::string d(dir);
  Reader r;
  if (!slurp(d + "/inputs.dat", r.b)) return fail(MSC_EIO, "cannot read %s/inputs.dat", dir);
  std::memset(p, 0, sizeof *p);
  bool ok = r.get(&p->dopml, 4) && r.get(&p->savefields, 4) && r.get(&p->f0, 8) && r.get(&p->fsig, 8) &&
            r.get(&p->sourcetype, 4) && r.get(&p->sourcepol, 4) && r.get(&p->gridfactor, 8) && r.get(&p->dt, 8) &&
            r.get(&p->tsteps, 4) && r.get(&p->nx, 4) && r.get(&p->ny, 4) && r.get(&p->nz, 4) &&
            r.get(p->metsize, 16) && r.get(&p->wpmax, 8) && r.get(&p->numax, 8) && r.get(p->metloc, 24) &&
            r.get(&p->metangle, 8) && r.get(&p->mettype, 4) && r.get(&p->numfiles, 4) && r.get(&p->nprobes, 4);
  if (!ok) return fail(MSC_EIO, "inputs.dat truncated before the probe list");
  if (p->nprobes < 0 || p->nprobes > MSC_MAX_PROBES)
    return fail(MSC_EINVAL, "inputs.dat: nprobes %d out of range", p->nprobes);
  ok = r.get(p->probex, 4 * (size_t)p->nprobes) && r.get(p->probey, 4 * (size_t)p->nprobes) &&
       r.get(p->probez, 4 * (size_t)p->nprobes) && r.get(&p->nntf, 4);
  if (!ok) return fail(MSC_EIO, "inputs.dat truncated in the probe list");
  if (p->nntf < 1 || p->nntf > MSC_MAX_NTF) return fail(MSC_EINVAL, "inputs.dat: nntf %d out of range", p->nntf);
  ok = r.get(p->fntf, 8 * (size_t)p->nntf) && r.get(&p->doplate, 4) && r.get(&p->platesize, 8);
  if (!ok) return fail(MSC_EIO, "inputs.dat truncated in the NTF list / plate fields");
  // doB0: read by the fork (meteor-fork.cpp:115) but never written by
  // setup/meteorrun.m; absent means Young's DI current (0).
  p->doB0 = 0;
  if (r.off + 4 <= r.b.size()) {
    r.get(&p->doB0, 4);
    if (r.off + 4 <= r.b.size() && std::memcmp(r.b.data() + r.off, kExtMagic, 4) == 0) {
      r.off += 4;
      int32_t ver = 0;
      ok = r.get(&ver, 4);
      if (!ok || ver < 1 || ver > kExtVersion)
        return fail(MSC_EIO, "inputs.dat: unknown extension record version %d", ver);
      ok = r.get(&p->pmllen, 4) && r.get(&p->media3d, 4) && r.get(&p->nu_scale_height, 8) && r.get(&p->src_tau, 8);
      if (ok && ver >= 2) ok = r.get(&p->j_collocated, 4);
      if (!ok) return fail(MSC_EIO, "inputs.dat: truncated extension record");
    }
  }
  std::vector<unsigned char> bb;
  if (!slurp(d + "/B0.dat", bb)) {
    if (p->doB0) return fail(MSC_EIO, "cannot read %s/B0.dat (doB0=1)", dir);
  } else {
    if (bb.size() < 24) return fail(MSC_EIO, "B0.dat truncated (%zu bytes)", bb.size());
    std::memcpy(p->B0, bb.data(), 24);
  }
  int rc = msc_params_validate(p);
  if (rc) return rc;
  size_t n = (size_t)p->nx * p->ny * (p->media3d ? (size_t)p->nz : 1);
  double* a = (double*)std::malloc(n * 8);
  double* c = (double*)std::malloc(n * 8);
  if (!a || !c) {
    std::free(a);
    std::free(c);
    return fail(MSC_ENOMEM, "host allocation of media maps failed");
  }
  const char* names[2] = {"/wp.dat", "/nu.dat"};
  double* dst[2] = {a, c};
  for (int q = 0; q < 2; ++q) {
    FILE* f = std::fopen((d + names[q]).c_str(), "rb");
    size_t got = 0;
    if (f) {
      got = std::fread(dst[q], 8, n, f);
      std::fclose(f);
    }
    if (got != n) {
      std::free(a);
      std::free(c);
      return fail(MSC_EIO, "%s%s: expected %zu doubles, read %zu", dir, names[q], n, got);
    }
  }
  *wp = a;
  *nu = c;
  return MSC_OK;
}

int msc_write_run_dir(const char* dir, const msc_params* p, const double* wp, const double* nu) {
  int rc = msc_params_validate(p);
  if (rc) return rc;
  std::string d(dir);
  std::vector<unsigned char> b;
  auto put = [&](const void* s, size_t n) {
    const unsigned char* c = (const unsigned char*)s;
    b.insert(b.end(), c, c + n);
  };
  put(&p->dopml, 4); put(&p->savefields, 4); put(&p->f0, 8); put(&p->fsig, 8);
  put(&p->sourcetype, 4); put(&p->sourcepol, 4); put(&p->gridfactor, 8); put(&p->dt, 8);
  put(&p->tsteps, 4); put(&p->nx, 4); put(&p->ny, 4); put(&p->nz, 4);
  put(p->metsize, 16); put(&p->wpmax, 8); put(&p->numax, 8); put(p->metloc, 24);
  put(&p->metangle, 8); put(&p->mettype, 4); put(&p->numfiles, 4); put(&p->nprobes, 4);
  put(p->probex, 4 * (size_t)p->nprobes); put(p->probey, 4 * (size_t)p->nprobes); put(p->probez, 4 * (size_t)p->nprobes);
  put(&p->nntf, 4); put(p->fntf, 8 * (size_t)p->nntf); put(&p->doplate, 4); put(&p->platesize, 8);
  put(&p->doB0, 4);
  bool ext = !((p->pmllen == 0 || p->pmllen == 10) && !p->media3d && p->nu_scale_height == 0 && p->src_tau == 0 &&
               !p->j_collocated);
  if (ext) {
    put(kExtMagic, 4);
    put(&kExtVersion, 4);
    put(&p->pmllen, 4); put(&p->media3d, 4); put(&p->nu_scale_height, 8); put(&p->src_tau, 8);
    put(&p->j_collocated, 4);
  }
  if (!spit(d + "/inputs.dat", b.data(), b.size())) return fail(MSC_EIO, "cannot write %s/inputs.dat", dir);
  if (!spit(d + "/B0.dat", p->B0, 24)) return fail(MSC_EIO, "cannot write %s/B0.dat", dir);
  size_t n = (size_t)p->nx * p->ny * (p->media3d ? (size_t)p->nz : 1);
  if (!spit(d + "/wp.dat", wp, n * 8) || !spit(d + "/nu.dat", nu, n * 8))
    return fail(MSC_EIO, "cannot write %s/wp.dat or nu.dat", dir);
  return MSC_OK;
}

}  // extern "C"
