// On-disk formats of the reference around the propagation path (host code):
//   wf_bar.bin / wf_bar.txt   ground-state orbitals handed to TDDFT
//                             (writewf, src/write_wf.f90:24-118; read_wf_tddft,
//                             src/read_wf.f90:130-228)
//   eig_bar                   eigenvalues (src/diag.f90:109-115)
//   d_dip.dat / dip.dat       dipole series (src/tddft_plot.f90:100-105)
//   tddft_save/               restart: info, phi_bar.bin, phi_bar_pert.bin
//                             (src/tddft_write_to_restart.f90,
//                             src/read_phi_bar.f90)
//
// Binary files are Fortran unformatted sequential files: every WRITE is one
// record framed by its byte count as a 4-byte integer before and after the
// payload; default integers are 4 bytes and reals 8.  Text files are written
// so that Fortran list-directed READ reads them (whitespace-separated values,
// one WRITE per line); d_dip.dat uses the reference's explicit format
// (' ',4e18.8).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <new>
#include <stdexcept>
#include <vector>

#include "opaw_hip.h"

namespace opaw {
int io_fail(int code, const char* fmt, ...);   // sets opaw_last_error (opaw_ctx.hip)
int ctx_orbitals_to_host(opaw_ctx* c, int sys, std::vector<std::complex<double>>& out, int* nb, long long* n);
int ctx_orbitals_from_host(opaw_ctx* c, int sys, const std::vector<std::complex<double>>& in);
int ctx_shape(opaw_ctx* c, int* nsys, int* nb, long long* n);
}  // namespace opaw

namespace {

using opaw::io_fail;

// ---------------------------------------------------------------------------
// Fortran unformatted sequential records
struct RecWriter {
  FILE* f = nullptr;
  std::vector<unsigned char> buf;
  void put(const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    buf.insert(buf.end(), c, c + n);
  }
  void tag(const char* t) {   // character*8, blank padded
    char s[8];
    std::memset(s, ' ', 8);
    std::memcpy(s, t, std::min<size_t>(8, std::strlen(t)));
    put(s, 8);
  }
  void i4(int v) { put(&v, 4); }
  void r8(double v) { put(&v, 8); }
  bool end() {
    if (buf.size() > 0x7fffffffu) return false;
    const int n = (int)buf.size();
    const bool ok = std::fwrite(&n, 4, 1, f) == 1 && (n == 0 || std::fwrite(buf.data(), 1, buf.size(), f) == buf.size()) &&
                    std::fwrite(&n, 4, 1, f) == 1;
    buf.clear();
    return ok;
  }
};

struct RecReader {
  FILE* f = nullptr;
  std::vector<unsigned char> rec;
  size_t pos = 0;
  bool next() {
    int n = 0, m = 0;
    if (std::fread(&n, 4, 1, f) != 1 || n < 0) return false;
    rec.resize((size_t)n);
    if (n > 0 && std::fread(rec.data(), 1, (size_t)n, f) != (size_t)n) return false;
    if (std::fread(&m, 4, 1, f) != 1 || m != n) return false;
    pos = 0;
    return true;
  }
  bool get(void* p, size_t n) {
    if (pos + n > rec.size()) return false;
    std::memcpy(p, rec.data() + pos, n);
    pos += n;
    return true;
  }
  bool tag(std::string& t) {
    char s[9] = {0};
    if (!get(s, 8)) return false;
    t.assign(s, 8);
    while (!t.empty() && t.back() == ' ') t.pop_back();
    return true;
  }
  bool i4(int& v) { return get(&v, 4); }
  bool r8(double& v) { return get(&v, 8); }
};

// ---------------------------------------------------------------------------
// list-directed text: whitespace-separated tokens (commas and slashes of
// Fortran list-directed input are not produced by the reference's writes)
struct TokReader {
  FILE* f = nullptr;
  bool next(std::string& s) {
    s.clear();
    int ch;
    while ((ch = std::fgetc(f)) != EOF && (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == ','))
      ;
    if (ch == EOF) return false;
    bool quoted = ch == '\'' || ch == '"';
    const int q = ch;
    if (!quoted) s.push_back((char)ch);
    while ((ch = std::fgetc(f)) != EOF) {
      if (quoted) {
        if (ch == q) break;
      } else if (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == ',') {
        break;
      }
      s.push_back((char)ch);
    }
    return true;
  }
  bool i(int& v) {
    std::string s;
    if (!next(s)) return false;
    char* e = nullptr;
    v = (int)std::strtol(s.c_str(), &e, 10);
    return e && *e == 0;
  }
  bool r(double& v) {
    std::string s;
    if (!next(s)) return false;
    for (char& ch : s)
      if (ch == 'd' || ch == 'D') ch = 'E';   // Fortran double exponent letter
    char* e = nullptr;
    v = std::strtod(s.c_str(), &e);
    return e && *e == 0;
  }
};

void put_real(FILE* f, double v) { std::fprintf(f, " %24.16E", v); }

// Fortran Ew.d edit descriptor with a zero leading digit (0.dddE+ee),
// right-justified in w columns
std::string fortran_e(double v, int w, int d) {
  char buf[64];
  std::string body;
  if (v == 0.0) {
    body = (std::signbit(v) ? "-0." : "0.") + std::string(d, '0') + "E+00";   // -0.0 keeps its sign
  } else {
    std::snprintf(buf, sizeof buf, "%.*E", d - 1, std::fabs(v));   // m.mmmE+xx with d digits
    std::string s(buf);
    const size_t epos = s.find('E');
    std::string digits = s.substr(0, 1) + s.substr(2, epos - 2);
    int ex = std::atoi(s.c_str() + epos + 1) + 1;
    char eb[16];
    if (std::abs(ex) <= 99)
      std::snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
    else
      std::snprintf(eb, sizeof eb, "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
    body = "0." + digits + eb;
    if (v < 0) body = "-" + body;
  }
  if ((int)body.size() >= w) return body.size() == (size_t)w ? body : std::string(w, '*');
  return std::string(w - body.size(), ' ') + body;
}

bool same_real(double a, double b) { return std::fabs(a - b) <= 1e-10 * std::max(1.0, std::fabs(a)); }

}  // namespace

extern "C" {

int opaw_write_wf(const char* path, int binary, int short_header, int nx, int ny, int nz, double dx, double dy,
                  double dz, int nstates, int ihomo, const double* eig, const void* phi) try {
  if (!path || !eig || !phi) return io_fail(OPAW_ERR_ARG, "null argument");
  if (nx <= 0 || ny <= 0 || nz <= 0 || nstates <= 0) return io_fail(OPAW_ERR_ARG, "bad shape");
  const long long N = (long long)nx * ny * nz;
  const std::complex<double>* ph = static_cast<const std::complex<double>*>(phi);
  FILE* f = std::fopen(path, binary ? "wb" : "w");
  if (!f) return io_fail(OPAW_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
  // the stored orbital: sign(Re phi) |phi| (src/write_wf.f90:55-63)
  std::vector<double> tmp((size_t)N);
  auto fold = [&](int i) {
    for (long long r = 0; r < N; ++r) {
      const std::complex<double> v = ph[(size_t)i * N + r];
      const double m = std::hypot(v.real(), v.imag());   // |phi| as libm's cabs
      tmp[(size_t)r] = v.real() > 0.0 ? m : -m;
    }
  };
  bool ok = true;
  if (binary) {
    RecWriter w{f};
    const char* it[] = {"nx", "ny", "nz"};
    const int iv[] = {nx, ny, nz};
    for (int k = 0; k < 3; ++k) {
      w.tag(it[k]);
      w.i4(iv[k]);
      ok = ok && w.end();
    }
    const char* rt[] = {"dx", "dy", "dz"};
    const double rv[] = {dx, dy, dz};
    for (int k = 0; k < 3; ++k) {
      w.tag(rt[k]);
      w.r8(rv[k]);
      ok = ok && w.end();
    }
    // writewf adds orbmn_wf and ihomo; the short header of
    // write_wf_for_sgw_debug (src/write_wf.f90:146-153) is what
    // read_wf_tddft reads
    const char* jt[] = {"nsp", "nstates", "orbmn_wf", "ihomo"};
    const int jv[] = {1, nstates, 1, ihomo};
    for (int k = 0; k < (short_header ? 2 : 4); ++k) {
      w.tag(jt[k]);
      w.i4(jv[k]);
      ok = ok && w.end();
    }
    w.tag("evls");
    ok = ok && w.end();
    for (int i = 0; i < nstates; ++i) w.r8(eig[i]);
    ok = ok && w.end();
    w.tag("orbitals");
    ok = ok && w.end();
    for (int i = 0; i < nstates && ok; ++i) {
      w.i4(i + 1);
      w.i4(1);
      ok = ok && w.end();
      fold(i);
      w.put(tmp.data(), tmp.size() * sizeof(double));
      ok = ok && w.end();
    }
  } else {
    std::fprintf(f, " nx       %d\n ny       %d\n nz       %d\n", nx, ny, nz);
    std::fprintf(f, " dx      ");
    put_real(f, dx);
    std::fprintf(f, "\n dy      ");
    put_real(f, dy);
    std::fprintf(f, "\n dz      ");
    put_real(f, dz);
    if (short_header)
      std::fprintf(f, "\n nsp      1\n nstates  %d\n evls\n", nstates);
    else
      std::fprintf(f, "\n nsp      1\n nstates  %d\n orbmn_wf 1\n ihomo    %d\n evls\n", nstates, ihomo);
    for (int i = 0; i < nstates; ++i) put_real(f, eig[i]);
    std::fprintf(f, "\n orbitals\n");
    for (int i = 0; i < nstates; ++i) {
      std::fprintf(f, " %d 1\n", i + 1);
      fold(i);
      for (long long r = 0; r < N; ++r) put_real(f, tmp[(size_t)r]);
      std::fprintf(f, "\n");
    }
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? OPAW_OK : io_fail(OPAW_ERR_ARG, "write error on %s", path);
} catch (const std::bad_alloc&) {
  return io_fail(OPAW_ERR_ARG, "%s: out of host memory", "opaw_write_wf");
} catch (const std::exception& e) {
  return io_fail(OPAW_ERR_ARG, "%s: %s", "opaw_write_wf", e.what());
}

int opaw_read_wf(const char* path, int binary, int nx, int ny, int nz, double dx, double dy, double dz,
                 int max_states, int* nstates, int* ihomo, double* eig, void* phi) try {
  if (!path || !nstates) return io_fail(OPAW_ERR_ARG, "null argument");
  FILE* f = std::fopen(path, binary ? "rb" : "r");
  if (!f) return io_fail(OPAW_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
  int hv[3] = {0, 0, 0}, nsp = 0, ns = 0, orbmn = 0, ih = 0;
  double dv[3] = {0, 0, 0};
  std::vector<double> ev;
  const long long N = (long long)nx * ny * nz;
  std::vector<double> tmp((size_t)std::max(0LL, N));
  auto bad = [&](const char* what) {
    std::fclose(f);
    return io_fail(OPAW_ERR_ARG, "%s: %s", path, what);
  };
  if (binary) {
    RecReader r{f};
    std::string t;
    for (int k = 0; k < 3; ++k)
      if (!r.next() || !r.tag(t) || !r.i4(hv[k])) return bad("bad grid header");
    for (int k = 0; k < 3; ++k)
      if (!r.next() || !r.tag(t) || !r.r8(dv[k])) return bad("bad spacing header");
    if (!r.next() || !r.tag(t) || !r.i4(nsp)) return bad("bad nsp");
    if (!r.next() || !r.tag(t) || !r.i4(ns)) return bad("bad nstates");
    if (ns < 0 || ns > max_states) return bad("nstates outside 0..max_states");
    // the reference's reader stops its header at nstates and skips the
    // eigenvalue tag, the eigenvalues and the orbital tag as three records;
    // files of writewf hold orbmn_wf and ihomo first
    if (!r.next() || !r.tag(t)) return bad("bad header");
    if (t == "orbmn_wf") {
      if (!r.i4(orbmn)) return bad("bad orbmn_wf");
      if (!r.next() || !r.tag(t) || !r.i4(ih)) return bad("bad ihomo");
      if (!r.next() || !r.tag(t)) return bad("missing evls");
    }
    if (!r.next()) return bad("missing eigenvalues");
    ev.resize((size_t)ns);
    for (int i = 0; i < ns; ++i)
      if (!r.r8(ev[(size_t)i])) return bad("short eigenvalue record");
    if (!r.next() || !r.tag(t)) return bad("missing orbitals tag");
  } else {
    TokReader r{f};
    std::string t;
    for (int k = 0; k < 3; ++k)
      if (!r.next(t) || !r.i(hv[k])) return bad("bad grid header");
    for (int k = 0; k < 3; ++k)
      if (!r.next(t) || !r.r(dv[k])) return bad("bad spacing header");
    if (!r.next(t) || !r.i(nsp)) return bad("bad nsp");
    if (!r.next(t) || !r.i(ns)) return bad("bad nstates");
    if (ns < 0 || ns > max_states) return bad("nstates outside 0..max_states");
    if (!r.next(t)) return bad("bad header");
    if (t == "orbmn_wf") {
      if (!r.i(orbmn) || !r.next(t) || !r.i(ih) || !r.next(t)) return bad("bad ihomo");
    }
    ev.resize((size_t)ns);
    for (int i = 0; i < ns; ++i)
      if (!r.r(ev[(size_t)i])) return bad("short eigenvalue list");
    if (!r.next(t)) return bad("missing orbitals tag");
  }
  // the reference's checks (src/read_wf.f90:181-187)
  if (hv[0] != nx || hv[1] != ny || hv[2] != nz) return bad("grid differs from the calculation (nx, nx_r ...)");
  if (!same_real(dx, dv[0]) || !same_real(dy, dv[1]) || !same_real(dz, dv[2])) return bad("spacing differs (dx, dx_r ...)");
  if (nsp != 1) return bad("nsp, nsp_r: only one spin channel");
  if (ns > max_states) return bad("more states in the file than max_states");
  std::complex<double>* out = static_cast<std::complex<double>*>(phi);
  for (int i = 0; i < ns; ++i) {
    int io = 0, sp = 0;
    if (binary) {
      RecReader r{f};
      if (!r.next() || !r.i4(io) || !r.i4(sp)) return bad("bad state header");
      if (!r.next() || r.rec.size() != (size_t)N * 8) return bad("bad state record");
      std::memcpy(tmp.data(), r.rec.data(), (size_t)N * 8);
    } else {
      TokReader r{f};
      if (!r.i(io) || !r.i(sp)) return bad("bad state header");
      for (long long k = 0; k < N; ++k)
        if (!r.r(tmp[(size_t)k])) return bad("short state");
    }
    if (out)
      for (long long k = 0; k < N; ++k) out[(size_t)i * N + k] = std::complex<double>(tmp[(size_t)k], 0.0);
  }
  std::fclose(f);
  *nstates = ns;
  if (ihomo) *ihomo = ih;
  if (eig)
    for (int i = 0; i < ns; ++i) eig[i] = ev[(size_t)i];
  return OPAW_OK;
} catch (const std::bad_alloc&) {
  return io_fail(OPAW_ERR_ARG, "%s: out of host memory", "opaw_read_wf");
} catch (const std::exception& e) {
  return io_fail(OPAW_ERR_ARG, "%s: %s", "opaw_read_wf", e.what());
}

int opaw_write_eig(const char* path, int n, const double* eig) try {
  if (!path || !eig || n <= 0) return io_fail(OPAW_ERR_ARG, "bad argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return io_fail(OPAW_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
  for (int i = 0; i < n; ++i) put_real(f, eig[i]);
  std::fprintf(f, "\n");
  return std::fclose(f) == 0 ? OPAW_OK : io_fail(OPAW_ERR_ARG, "write error on %s", path);
} catch (const std::bad_alloc&) {
  return io_fail(OPAW_ERR_ARG, "%s: out of host memory", "opaw_write_eig");
} catch (const std::exception& e) {
  return io_fail(OPAW_ERR_ARG, "%s: %s", "opaw_write_eig", e.what());
}

int opaw_read_eig(const char* path, int n, double* eig) try {
  if (!path || !eig || n <= 0) return io_fail(OPAW_ERR_ARG, "bad argument");
  FILE* f = std::fopen(path, "r");
  if (!f) return io_fail(OPAW_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
  TokReader r{f};
  for (int i = 0; i < n; ++i)
    if (!r.r(eig[i])) {
      std::fclose(f);
      return io_fail(OPAW_ERR_ARG, "%s: fewer than %d eigenvalues", path, n);
    }
  std::fclose(f);
  return OPAW_OK;
} catch (const std::bad_alloc&) {
  return io_fail(OPAW_ERR_ARG, "%s: out of host memory", "opaw_read_eig");
} catch (const std::exception& e) {
  return io_fail(OPAW_ERR_ARG, "%s: %s", "opaw_read_eig", e.what());
}

int opaw_write_dipoles(const char* d_dip_path, const char* dip_path, int append, int it0, int nt, double dt,
                       const double* d_dip, const double* dip, int nsys) try {
  if (nt <= 0 || nsys <= 0) return io_fail(OPAW_ERR_ARG, "bad argument");
  const char* mode = append ? "a" : "w";
  if (d_dip_path && d_dip) {
    FILE* f = std::fopen(d_dip_path, mode);
    if (!f) return io_fail(OPAW_ERR_ARG, "cannot open %s: %s", d_dip_path, std::strerror(errno));
    for (int k = 0; k < nt; ++k) {
      const double t = (double)(it0 + k) * dt;   // (it-1)*dt with it = it0+k+1
      std::string line = " " + fortran_e(t, 18, 8);
      for (int q = 0; q < 3; ++q) line += fortran_e(d_dip[(size_t)3 * k + q], 18, 8);
      std::fprintf(f, "%s\n", line.c_str());
    }
    if (std::fclose(f) != 0) return io_fail(OPAW_ERR_ARG, "write error on %s", d_dip_path);
  }
  if (dip_path && dip) {
    FILE* f = std::fopen(dip_path, mode);
    if (!f) return io_fail(OPAW_ERR_ARG, "cannot open %s: %s", dip_path, std::strerror(errno));
    for (int k = 0; k < nt; ++k) {
      put_real(f, (double)(it0 + k) * dt);
      for (int q = 0; q < 3 * nsys; ++q) put_real(f, dip[(size_t)3 * nsys * k + q]);
      std::fprintf(f, "\n");
    }
    if (std::fclose(f) != 0) return io_fail(OPAW_ERR_ARG, "write error on %s", dip_path);
  }
  return OPAW_OK;
} catch (const std::bad_alloc&) {
  return io_fail(OPAW_ERR_ARG, "%s: out of host memory", "opaw_write_dipoles");
} catch (const std::exception& e) {
  return io_fail(OPAW_ERR_ARG, "%s: %s", "opaw_write_dipoles", e.what());
}

// restart files: system 0 -> phi_bar.bin, 1 -> phi_bar_pert.bin,
// s >= 2 -> phi_bar_pert<s>.bin (further kick directions)
static std::string restart_file(const char* dir, int s) {
  std::string d(dir);
  if (s == 0) return d + "/phi_bar.bin";
  if (s == 1) return d + "/phi_bar_pert.bin";
  return d + "/phi_bar_pert" + std::to_string(s) + ".bin";
}

int opaw_write_restart(opaw_ctx* ctx, const char* dir, int it) try {
  if (!ctx || !dir) return io_fail(OPAW_ERR_ARG, "null argument");
  int nsys = 0, nb = 0;
  long long n = 0;
  int r = opaw::ctx_shape(ctx, &nsys, &nb, &n);
  if (r) return r;
  mkdir(dir, 0755);   // an existing directory is fine (src/tddft_write_to_restart.f90:22-28)
  {
    const std::string info = std::string(dir) + "/info";
    FILE* f = std::fopen(info.c_str(), "w");
    if (!f) return io_fail(OPAW_ERR_ARG, "cannot open %s: %s", info.c_str(), std::strerror(errno));
    std::fprintf(f, " %d %d %d\n", nb, 1, it);   // nb, nds (one process), it
    if (std::fclose(f) != 0) return io_fail(OPAW_ERR_ARG, "write error on %s", info.c_str());
  }
  std::vector<std::complex<double>> host;
  for (int s = 0; s < nsys; ++s) {
    int nb2 = 0;
    long long n2 = 0;
    if ((r = opaw::ctx_orbitals_to_host(ctx, s, host, &nb2, &n2))) return r;
    const std::string fn = restart_file(dir, s);
    FILE* f = std::fopen(fn.c_str(), "wb");
    if (!f) return io_fail(OPAW_ERR_ARG, "cannot open %s: %s", fn.c_str(), std::strerror(errno));
    RecWriter w{f};
    bool ok = true;
    for (int i = 0; i < nb2 && ok; ++i) {   // one record per orbital, complex*16
      w.put(host.data() + (size_t)i * n2, (size_t)n2 * sizeof(std::complex<double>));
      ok = w.end();
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return io_fail(OPAW_ERR_ARG, "write error on %s", fn.c_str());
  }
  return OPAW_OK;
} catch (const std::bad_alloc&) {
  return io_fail(OPAW_ERR_ARG, "%s: out of host memory", "opaw_write_restart");
} catch (const std::exception& e) {
  return io_fail(OPAW_ERR_ARG, "%s: %s", "opaw_write_restart", e.what());
}

int opaw_read_restart(opaw_ctx* ctx, const char* dir, int* it) try {
  if (!ctx || !dir) return io_fail(OPAW_ERR_ARG, "null argument");
  int nsys = 0, nb = 0;
  long long n = 0;
  int r = opaw::ctx_shape(ctx, &nsys, &nb, &n);
  if (r) return r;
  int nb_r = 0, nds_r = 0, it_r = 0;
  {
    const std::string info = std::string(dir) + "/info";
    FILE* f = std::fopen(info.c_str(), "r");
    if (!f) return io_fail(OPAW_ERR_ARG, "cannot open %s: %s", info.c_str(), std::strerror(errno));
    TokReader t{f};
    const bool ok = t.i(nb_r) && t.i(nds_r) && t.i(it_r);
    std::fclose(f);
    if (!ok) return io_fail(OPAW_ERR_ARG, "%s: bad info record", info.c_str());
  }
  // src/read_phi_bar.f90:13-14: one process group here
  if (nds_r != 1) return io_fail(OPAW_ERR_ARG, "nds_read/=nds (%d process groups in the restart)", nds_r);
  if (nb_r != nb) return io_fail(OPAW_ERR_ARG, "restart has %d orbitals, the context %d", nb_r, nb);
  std::vector<std::complex<double>> host((size_t)nb * n);
  for (int s = 0; s < nsys; ++s) {
    const std::string fn = restart_file(dir, s);
    FILE* f = std::fopen(fn.c_str(), "rb");
    if (!f) return io_fail(OPAW_ERR_ARG, "cannot open %s: %s", fn.c_str(), std::strerror(errno));
    RecReader rr{f};
    for (int i = 0; i < nb; ++i) {
      if (!rr.next() || rr.rec.size() != (size_t)n * sizeof(std::complex<double>)) {
        std::fclose(f);
        return io_fail(OPAW_ERR_ARG, "%s: bad orbital record %d", fn.c_str(), i + 1);
      }
      std::memcpy(host.data() + (size_t)i * n, rr.rec.data(), rr.rec.size());
    }
    std::fclose(f);
    if ((r = opaw::ctx_orbitals_from_host(ctx, s, host))) return r;
  }
  if (it) *it = it_r;
  return OPAW_OK;
} catch (const std::bad_alloc&) {
  return io_fail(OPAW_ERR_ARG, "%s: out of host memory", "opaw_read_restart");
} catch (const std::exception& e) {
  return io_fail(OPAW_ERR_ARG, "%s: %s", "opaw_read_restart", e.what());
}

}  // extern "C"
