// egress.cpp -- visibility egress on the host: the client's CSV writer
// (MDT_OneStation_client.cu:1202-1272) and a self-describing binary dump.
//
// CSV rows, as the reference writes them: for element j, pol p, element
// k >= j, pol q >= p, channel z = 0..N-1 in fftshift order (source channel
// N/2..N-1 then 0..N/2-1), the line "j,p,k,q,z,re,im" with "%ld,%d,%ld,%d,
// %ld,%.6f,%.6f".  The value is V[a][b] = alpha sum_s x_a conj(x_b) with
// a = j*Npols + p, b = k*Npols + q (a <= b): the reference reads it from the
// column-major cuBLAS output XOutput[c*NN*NN + a + b*NN].  Starting q at p for
// every k skips the (j, 1, k, 0) cross-polarisation rows when k > j; with
// all_pol_pairs = 1 they are written too (q from 0 when k > j).
#include <cstdio>
#include <cstring>
#include <ctime>
#include <vector>

#include "mdtfx.h"

namespace mdt {
int fail(int code, const char *fmt, ...);
}
using mdt::fail;

namespace {

// V[a][b], a <= b, from either layout
inline void vis_at(const mdt_complex *vis, int layout, long c, int NN, int a, int b, float *re, float *im) {
  if (layout == MDT_VIS_FULL_COLMAJOR) {
    const mdt_complex v = vis[c * (long)NN * NN + a + (long)b * NN];
    *re = v.x;
    *im = v.y;
  } else {  // packed lower triangle holds V[b][a] (b >= a) = conj(V[a][b]); the diagonal is itself
    const mdt_complex v = vis[c * ((long)NN * (NN + 1) / 2) + (long)b * (b + 1) / 2 + a];
    *re = v.x;
    *im = a == b ? v.y : -v.y;
  }
}

}  // namespace

extern "C" {

int mdt_csv_filename(char *name, size_t size, long long unix_time) {
  if (!name) return fail(MDT_ERR_NULL_POINTER, "name is NULL");
  const time_t t = (time_t)unix_time;
  struct tm tmv;
  if (!localtime_r(&t, &tmv)) return fail(MDT_ERR_INVALID_VALUE, "time %lld", unix_time);
  if (strftime(name, size, "MDT_%Y%m%d_%H%M%S.csv", &tmv) == 0)
    return fail(MDT_ERR_INVALID_VALUE, "file name buffer of %zu bytes too small", size);
  return MDT_SUCCESS;
}

int mdt_write_csv(const char *path, const mdt_complex *vis, int layout, int Nelements, int Npols,
                  int Nchannels, int all_pol_pairs) {
  if (!path || !vis) return fail(MDT_ERR_NULL_POINTER, "path or visibilities NULL");
  if (layout != MDT_VIS_FULL_COLMAJOR && layout != MDT_VIS_TRIANGLE)
    return fail(MDT_ERR_INVALID_VALUE, "unknown visibility layout %d", layout);
  if (Nelements < 1 || Npols < 1 || Npols > 2 || Nchannels < 2 || (Nchannels & 1))
    return fail(MDT_ERR_INVALID_VALUE, "shape Nelements=%d Npols=%d Nchannels=%d", Nelements, Npols,
                Nchannels);
  FILE *f = fopen(path, "w");
  if (!f) return fail(MDT_ERR_INVALID_VALUE, "cannot open or create %s", path);
  const int NN = Nelements * Npols;
  std::vector<char> out;
  out.reserve(1 << 22);
  char line[128];
  int st = MDT_SUCCESS;
  for (long j = 0; j < Nelements && st == MDT_SUCCESS; j++)
    for (int p = 0; p < Npols; p++)
      for (long k = j; k < Nelements; k++)
        for (int q = (all_pol_pairs && k > j) ? 0 : p; q < Npols; q++) {
          const int a = (int)j * Npols + p, b = (int)k * Npols + q;
          long z = 0;
          for (int half = 0; half < 2; half++) {
            const int c0 = half == 0 ? Nchannels / 2 : 0, c1 = half == 0 ? Nchannels : Nchannels / 2;
            for (int c = c0; c < c1; c++, z++) {
              float re, im;
              vis_at(vis, layout, c, NN, a, b, &re, &im);
              const int n = snprintf(line, sizeof line, "%ld,%d,%ld,%d,%ld,%.6f,%.6f\n", j, p, k, q, z,
                                     (double)re, (double)im);
              if (n <= 0 || n >= (int)sizeof line) {
                st = fail(MDT_ERR_INVALID_VALUE, "value too long to format at channel %d", c);
                break;
              }
              out.insert(out.end(), line, line + n);
            }
          }
          if (out.size() > (1u << 22)) {
            if (fwrite(out.data(), 1, out.size(), f) != out.size()) st = fail(MDT_ERR_INVALID_VALUE, "write to %s", path);
            out.clear();
          }
        }
  if (st == MDT_SUCCESS && !out.empty() && fwrite(out.data(), 1, out.size(), f) != out.size())
    st = fail(MDT_ERR_INVALID_VALUE, "write to %s", path);
  if (fclose(f) != 0 && st == MDT_SUCCESS) st = fail(MDT_ERR_INVALID_VALUE, "close %s", path);
  return st;
}

// binary dump: 64-byte header then the visibilities as they are in memory
//   char magic[8] = "MDTVIS01"; int32 layout, dtype (0 complex64, 1 complex128),
//   Nelements, Npols, Nchannels, Nspectra; int64 unix_time; 16 bytes zero
int mdt_write_binary(const char *path, const void *vis, int layout, int dtype, int Nelements, int Npols,
                     int Nchannels, int Nspectra, long long unix_time) {
  if (!path || !vis) return fail(MDT_ERR_NULL_POINTER, "path or visibilities NULL");
  if (layout != MDT_VIS_FULL_COLMAJOR && layout != MDT_VIS_TRIANGLE)
    return fail(MDT_ERR_INVALID_VALUE, "unknown visibility layout %d", layout);
  if (dtype != 0 && dtype != 1) return fail(MDT_ERR_INVALID_VALUE, "unknown dtype %d", dtype);
  if (Nelements < 1 || Npols < 1 || Npols > 2 || Nchannels < 1)
    return fail(MDT_ERR_INVALID_VALUE, "shape Nelements=%d Npols=%d Nchannels=%d", Nelements, Npols,
                Nchannels);
  const long NN = (long)Nelements * Npols;
  const long per = layout == MDT_VIS_FULL_COLMAJOR ? NN * NN : NN * (NN + 1) / 2;
  const size_t bytes = (size_t)per * Nchannels * (dtype == 0 ? 8 : 16);
  unsigned char hdr[64] = {0};
  memcpy(hdr, "MDTVIS01", 8);
  const int32_t h[6] = {layout, dtype, Nelements, Npols, Nchannels, Nspectra};
  memcpy(hdr + 8, h, sizeof h);
  const int64_t t = unix_time;
  memcpy(hdr + 32, &t, 8);
  FILE *f = fopen(path, "wb");
  if (!f) return fail(MDT_ERR_INVALID_VALUE, "cannot open or create %s", path);
  int st = MDT_SUCCESS;
  if (fwrite(hdr, 1, 64, f) != 64 || fwrite(vis, 1, bytes, f) != bytes)
    st = fail(MDT_ERR_INVALID_VALUE, "write to %s", path);
  if (fclose(f) != 0 && st == MDT_SUCCESS) st = fail(MDT_ERR_INVALID_VALUE, "close %s", path);
  return st;
}

}  // extern "C"
