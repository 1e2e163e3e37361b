// smaug_io.cpp -- VAC configuration files of the reference, host side.
//
//   .ini  ASCII configuration: one header of five lines, then one line per
//         cell (i fastest, ghost cells included) holding the ndim positions
//         and the NVAR fields (reference readasciivacconfig /
//         writeasciivacconfig, src/readwrite.c:975-1082,1085-1300);
//   .out  binary snapshot: 79-byte name, int32 it, f64 t, int32 ndim,
//         neqpar, nw, int32 n[ndim] (ghosts included), f64 eqpar[neqpar],
//         79-byte variable names, then ndim position blocks and NVAR field
//         blocks of prod(n) doubles each, i fastest, and one '\n'
//         (reference writevacconfig, src/readwrite.c:160-366).
//
// The reference writes 79 bytes straight from its name and variable-name
// buffers, whatever follows the terminating NUL; here the padding is NUL
// bytes so files are reproducible.  Everything else is byte for byte the
// reference's layout (native little-endian, no record markers).
#include <errno.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/smaug_hip.h"

namespace smaug {
int io_fail(int code, const char* fmt, ...);  // defined in smaug_host.hip
}

namespace {

const char* varnames(int ndim) {
  // src/readwrite.c:305-309
  return ndim == 3 ? "x y z rho mx my mz e bx by bz gamma eta g1 g2 g3" : "x y rho mx my e bx by  gamma eta g1 g2";
}

long long ncell(int ndim, const int* n) {
  long long N = 1;
  for (int d = 0; d < ndim; d++) N *= n[d];
  return N;
}

bool check_dims(int ndim, const int* n) {
  if (ndim != 2 && ndim != 3) return false;
  for (int d = 0; d < ndim; d++)
    if (n[d] < 1) return false;
  return true;
}

bool read_line(FILE* f, std::string& out) {
  out.clear();
  int c;
  while ((c = fgetc(f)) != EOF) {
    if (c == '\n') return true;
    out.push_back((char)c);
  }
  return !out.empty();
}

void pad79(FILE* f, const char* s) {
  char buf[79];
  memset(buf, 0, sizeof buf);
  const size_t L = s ? strlen(s) : 0;
  memcpy(buf, s, L < sizeof buf ? L : sizeof buf);
  fwrite(buf, 1, sizeof buf, f);
}

}  // namespace

extern "C" int smaug_ini_read(const char* path, int ndim, const int* n, double* w, double* pos, int* it,
                              double* t) {
  using smaug::io_fail;
  if (!path || !n || !w || !pos || !check_dims(ndim, n)) return io_fail(SMAUG_E_ARG, "smaug_ini_read: bad arguments");
  FILE* f = fopen(path, "r");
  if (!f) return io_fail(SMAUG_E_ARG, "smaug_ini_read: cannot open %s (%s)", path, strerror(errno));
  std::string line;
  int rc = SMAUG_OK;
  int hit = 0, hnd = 0, hneq = 0, hnw = 0;
  double ht = 0;
  int hn[3] = {1, 1, 1};
  const int nvar = 4 + 3 * ndim;
  const long long N = ncell(ndim, n);
  // header: name; it t ndim neqpar nw; n[ndim]; eqpar; varnames
  if (!read_line(f, line) || !read_line(f, line) ||
      sscanf(line.c_str(), "%d %lf %d %d %d", &hit, &ht, &hnd, &hneq, &hnw) != 5) {
    rc = io_fail(SMAUG_E_ARG, "smaug_ini_read: %s: malformed header line 2", path);
  } else if (hnd != ndim || hnw != nvar) {
    rc = io_fail(SMAUG_E_ARG, "smaug_ini_read: %s holds ndim=%d nw=%d, expected %d and %d", path, hnd, hnw, ndim, nvar);
  } else if (!read_line(f, line) ||
             sscanf(line.c_str(), "%d %d %d", &hn[0], &hn[1], &hn[2]) < ndim) {
    rc = io_fail(SMAUG_E_ARG, "smaug_ini_read: %s: malformed grid line", path);
  } else {
    for (int d = 0; d < ndim; d++)
      if (hn[d] != n[d]) rc = io_fail(SMAUG_E_ARG, "smaug_ini_read: %s grid dim %d is %d, expected %d", path, d, hn[d], n[d]);
  }
  if (rc == SMAUG_OK && (!read_line(f, line) || !read_line(f, line)))
    rc = io_fail(SMAUG_E_ARG, "smaug_ini_read: %s: truncated header", path);
  // cells, i fastest: positions then the fields
  for (long long c = 0; rc == SMAUG_OK && c < N; c++) {
    for (int d = 0; d < ndim; d++)
      if (fscanf(f, "%lf", &pos[(long long)d * N + c]) != 1) {
        rc = io_fail(SMAUG_E_ARG, "smaug_ini_read: %s: cell %lld: short row", path, c);
        break;
      }
    for (int v = 0; rc == SMAUG_OK && v < nvar; v++)
      if (fscanf(f, "%lf", &w[(long long)v * N + c]) != 1)
        rc = io_fail(SMAUG_E_ARG, "smaug_ini_read: %s: cell %lld: short row", path, c);
  }
  fclose(f);
  if (rc == SMAUG_OK) {
    if (it) *it = hit;
    if (t) *t = ht;
  }
  return rc;
}

extern "C" int smaug_ini_write(const char* path, const char* name, int ndim, const int* n, const double* w,
                               const double* pos, int it, double t, double gamma, double eta, const double* g) {
  using smaug::io_fail;
  if (!path || !n || !w || !pos || !g || !check_dims(ndim, n)) return io_fail(SMAUG_E_ARG, "smaug_ini_write: bad arguments");
  FILE* f = fopen(path, "w");
  if (!f) return io_fail(SMAUG_E_ARG, "smaug_ini_write: cannot open %s (%s)", path, strerror(errno));
  const int nvar = 4 + 3 * ndim;
  const long long N = ncell(ndim, n);
  // src/readwrite.c:1220-1266: header with %lg, cells with %lG (6 digits)
  fprintf(f, "%s\n", name ? name : "");
  fprintf(f, "%d %lg %d %d %d\n", it, t, ndim, ndim == 3 ? 7 : 6, nvar);
  if (ndim == 3) {
    fprintf(f, "%d %d %d\n", n[0], n[1], n[2]);
    fprintf(f, "%lg %lg %lg %lg %lg %d %d\n", gamma, eta, g[0], g[1], g[2], 0, 0);
  } else {
    fprintf(f, "%d %d\n", n[0], n[1]);
    fprintf(f, "%lg %lg %lg %lg %d %d\n", gamma, eta, g[0], g[1], 0, 0);
  }
  fprintf(f, "%s\n", varnames(ndim));
  for (long long c = 0; c < N; c++) {
    for (int d = 0; d < ndim; d++) fprintf(f, d == 0 ? "%lG" : " %lG", pos[(long long)d * N + c]);
    for (int v = 0; v < nvar; v++) fprintf(f, " %lG", w[(long long)v * N + c]);
    fputc('\n', f);
  }
  const bool ok = ferror(f) == 0;
  fclose(f);
  return ok ? SMAUG_OK : io_fail(SMAUG_E_ARG, "smaug_ini_write: write error on %s", path);
}

extern "C" int smaug_out_write(const char* path, const char* name, int ndim, const int* n, const double* w,
                               const double* pos, int it, double t, double gamma, double eta, const double* g) {
  using smaug::io_fail;
  if (!path || !n || !w || !pos || !g || !check_dims(ndim, n)) return io_fail(SMAUG_E_ARG, "smaug_out_write: bad arguments");
  FILE* f = fopen(path, "wb");
  if (!f) return io_fail(SMAUG_E_ARG, "smaug_out_write: cannot open %s (%s)", path, strerror(errno));
  const int nvar = 4 + 3 * ndim;
  const long long N = ncell(ndim, n);
  pad79(f, name);
  const int i_it = it;
  fwrite(&i_it, sizeof(int), 1, f);
  fwrite(&t, sizeof(double), 1, f);
  const int hdr[3] = {ndim, ndim == 3 ? 7 : 6, nvar};
  fwrite(hdr, sizeof(int), 3, f);
  fwrite(n, sizeof(int), (size_t)ndim, f);
  // eqpar: gamma, eta, g0, g1, (g2,) 0, 0 (src/readwrite.c:281-300)
  double eq[7] = {gamma, eta, g[0], g[1], 0, 0, 0};
  if (ndim == 3) eq[4] = g[2];
  fwrite(eq, sizeof(double), ndim == 3 ? 7 : 6, f);
  pad79(f, varnames(ndim));
  fwrite(pos, sizeof(double), (size_t)(ndim * N), f);
  fwrite(w, sizeof(double), (size_t)(nvar * N), f);
  fputc('\n', f);
  const bool ok = ferror(f) == 0;
  fclose(f);
  return ok ? SMAUG_OK : io_fail(SMAUG_E_ARG, "smaug_out_write: write error on %s", path);
}
