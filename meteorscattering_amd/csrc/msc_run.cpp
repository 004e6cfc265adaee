// msc_run_dir: the whole drop-in process of meteor-fork.cpp:42-1742 on top
// of the solver handle - read the run directory, solve, write
// output_{E,H,J}.dat (appended per frame, :1636-1662), output_NTF.dat
// (:1671-1714), Probe.dat (:1716-1730) and log.txt (:661-674, :1605-1632,
// :1732-1738).
#include <sys/resource.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "msc_common.h"

using namespace msc;

namespace {

double now_min() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count() / 60.0;
}

bool append(const std::string& path, const double* d, size_t n) {
  FILE* f = std::fopen(path.c_str(), "ab");
  if (!f) return false;
  size_t w = std::fwrite(d, 8, n, f);
  return std::fclose(f) == 0 && w == n;
}

}  // namespace

extern "C" int msc_run_dir(const char* dir, int precision) {
  if (!dir) return fail(MSC_EINVAL, "NULL directory");
  const double tstart = now_min();
  const std::string d(dir);
  msc_params p;
  double *wp = nullptr, *nu = nullptr;
  int rc = msc_read_run_dir(dir, &p, &wp, &nu);
  if (rc) {
    std::ofstream log(d + "/log.txt");
    log << "Error reading inputs: " << msc_last_error() << "\n";
    return rc;
  }
  msc_solver* s = nullptr;
  rc = msc_solver_create(&p, wp, nu, precision, &s);
  msc_free(wp);
  msc_free(nu);
  if (rc) {
    std::ofstream log(d + "/log.txt");
    log << "Error creating the solver: " << msc_last_error() << "\n";
    return rc;
  }
  const double dx = (kC / p.f0) / p.gridfactor;
  struct rusage ru;
  getrusage(RUSAGE_SELF, &ru);
  {
    std::ofstream log(d + "/log.txt");
    log << "------------------------------------------\n";
    log << "3D Meteor Scattering Simulation\n";
    log << "------------------------------------------\n";
    log << "dx = " << dx << "\n";
    log << "source type = " << p.sourcetype << "\n";
    log << "source polarization = " << p.sourcepol << "\n";
    log << "Time step = " << p.dt << "\n";
    log << "Total time steps = " << p.tsteps << "\n";
    log << "Total NTF frequencies = " << p.nntf << "\n";
    log << "last NTF frequency = " << p.fntf[p.nntf - 1] << "\n";
    log << "Total memory usage = " << (double)ru.ru_maxrss / 1024 / 1024 << " GB\n";
    log << "Device memory usage = " << (double)msc_solver_device_bytes(s) / (1 << 30) << " GB\n";
  }
  const int T = p.tsteps, cad = p.tsteps / p.numfiles;
  const size_t fn = msc_solver_frame_count(s);
  const size_t fsz = fn - 2 * (size_t)p.nx;
  std::vector<double> ef(fn), hf(fn), jf(p.savefields ? fsz : 0, 0.0);
  const double tloop = now_min();
  int t = 0;
  while (t < T) {
    // run to the next cadence step (or step 20 for the runtime projection)
    int next = (t / cad + 1) * cad;
    if (t <= 20 && next > 21) next = 21;
    if (next > T) next = T;
    if ((rc = msc_solver_step(s, next - t)) || (rc = msc_solver_sync(s))) break;
    for (int u = t; u < next; u++) {
      if (u == 20) {
        const double mins = now_min() - tloop;
        std::ofstream log(d + "/log.txt", std::ios::app);
        log << "t = 20, time taken = " << mins << " minutes.\n";
        log << "You can expect the total simulation to take " << mins * T / 20.0 << " minutes.\n";
      }
      if (u % cad) continue;
      const int f = u / cad;
      double emax = 0;
      rc = msc_solver_get_frame(s, f, p.savefields ? ef.data() : nullptr, p.savefields ? hf.data() : nullptr, &emax);
      if (rc) break;
      {
        std::ofstream log(d + "/log.txt", std::ios::app);
        const double partial = 100 * u / T;  // integer division as :1617
        log << "t = " << u << ": " << partial << " % Done...  ";
        log << "Maximum abs(Ex) is " << emax << "\n";
      }
      if (p.savefields) {
        if (!append(d + "/output_E.dat", ef.data(), fn) || !append(d + "/output_J.dat", jf.data(), fsz) ||
            !append(d + "/output_H.dat", hf.data(), fn)) {
          rc = fail(MSC_EIO, "cannot append frame files in %s", dir);
          break;
        }
      }
    }
    if (rc) break;
    t = next;
  }
  if (!rc) {
    std::vector<double> pay(msc_solver_ntff_count(s));
    rc = msc_solver_get_ntff(s, pay.data(), pay.size());
    if (!rc) {
      FILE* f = std::fopen((d + "/output_NTF.dat").c_str(), "wb");
      if (!f) rc = fail(MSC_EIO, "cannot write output_NTF.dat");
      else {
        int32_t hdr[2] = {p.nntf, pml_len(&p) + 5};
        bool ok = std::fwrite(hdr, 4, 2, f) == 2 && std::fwrite(p.fntf, 8, p.nntf, f) == (size_t)p.nntf &&
                  std::fwrite(pay.data(), 8, pay.size(), f) == pay.size();
        if (std::fclose(f) != 0 || !ok) rc = fail(MSC_EIO, "short write of output_NTF.dat");
      }
      // the reference prints My/Mz_front[39][79][0] (:1694-1695); only where that index exists
      const size_t M2 = 2 * (size_t)p.nntf;
      if (!rc && p.ny > 39 && p.nz > 79) {
        const size_t e = ((size_t)39 * p.nz + 79) * M2;
        const double* my = pay.data() + 2 * M2;
        const double* mz = my + (size_t)p.ny * p.nz * M2;
        std::cout << "My-front at y = 40, z = 80 is " << my[e] << "\n";
        std::cout << "Mz-front at y = 40, z = 80 is " << mz[e] << "\n";
      }
    }
  }
  if (!rc) {
    const size_t n = (size_t)6 * T * p.nprobes;
    std::vector<double> pr(n ? n : 1);
    rc = msc_solver_get_probes(s, pr.data(), n);
    if (!rc) {
      FILE* f = std::fopen((d + "/Probe.dat").c_str(), "wb");
      if (!f) rc = fail(MSC_EIO, "cannot write Probe.dat");
      else {
        bool ok = std::fwrite(&p.nprobes, 4, 1, f) == 1 && std::fwrite(p.probex, 4, p.nprobes, f) == (size_t)p.nprobes &&
                  std::fwrite(p.probey, 4, p.nprobes, f) == (size_t)p.nprobes &&
                  std::fwrite(p.probez, 4, p.nprobes, f) == (size_t)p.nprobes && std::fwrite(pr.data(), 8, n, f) == n;
        if (std::fclose(f) != 0 || !ok) rc = fail(MSC_EIO, "short write of Probe.dat");
      }
    }
  }
  msc_solver_destroy(s);
  std::ofstream log(d + "/log.txt", std::ios::app);
  if (rc) {
    log << "Error: " << msc_last_error() << "\n";
    return rc;
  }
  log << "All done!\n";
  log << "Total run time = " << now_min() - tstart << "minutes.\n";
  return MSC_OK;
}
