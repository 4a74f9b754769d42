/*
 * dense3d_oracle.c -- CPU restatement of write_dense3d_frame's serialization
 * (src/dense3d_sequence.rs:497-563).  TEST INFRASTRUCTURE ONLY (see
 * rheon_oracle.h): the checker of the device frame writer and its timed CPU
 * baseline; the product never links or calls it.
 *
 * Numbers: the reference prints f32 fields with `{v:.9e}` and f64 fields and
 * diagnostics with `{v:.17e}` (Rust core::fmt exact mode: the exact binary
 * value rounded to P + 1 significant digits, ties to even; exponent without
 * '+' or padding; '-' for negative values including -0.0).  Here glibc's
 * "%.*e" gives the same digits (it converts exactly and rounds in the
 * current rounding mode, to nearest-even by default), and the exponent is
 * rewritten to Rust's form.  The preflight checks are restated in Python
 * (oracle/__init__.py, dense3d_frame).
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rheon_oracle.h"

int or_format_sci(double v, int digits, char* out) {
  char tmp[64];
  int n = snprintf(tmp, sizeof tmp, "%.*e", digits - 1, v);
  char* e = strchr(tmp, 'e');
  if (n <= 0 || !e) return -1;
  int len = (int)(e - tmp);
  memcpy(out, tmp, (size_t)len);
  out[len++] = 'e';
  const char* p = e + 1;
  if (*p == '-') out[len++] = '-';
  if (*p == '-' || *p == '+') ++p;
  while (p[0] == '0' && p[1]) ++p; /* drop the exponent's leading zeros */
  while (*p) out[len++] = *p++;
  out[len] = 0;
  return len;
}

/* Appends text (or only counts it when out is NULL). */
static void put(char* out, int64_t* at, const char* text, int64_t len) {
  if (out) memcpy(out + *at, text, (size_t)len);
  *at += len;
}
static void puts_(char* out, int64_t* at, const char* text) { put(out, at, text, (int64_t)strlen(text)); }
static void put_sci(char* out, int64_t* at, double v, int digits) {
  char b[64];
  const int n = or_format_sci(v, digits, b);
  put(out, at, b, n);
}
static void put_u64(char* out, int64_t* at, uint64_t v) {
  char b[32];
  const int n = snprintf(b, sizeof b, "%" PRIu64, v);
  put(out, at, b, n);
}

/* The numbers of one field array, each preceded by ',' but the first. */
static int64_t put_array(char* out, int64_t at, const void* a, int f64, int64_t len, int threads) {
  const int digits = f64 ? 18 : 10;
  if (threads <= 1 || len < 65536) {
    for (int64_t i = 0; i < len; ++i) {
      if (i) puts_(out, &at, ",");
      put_sci(out, &at, f64 ? ((const double*)a)[i] : (double)((const float*)a)[i], digits);
    }
    return at;
  }
  /* chunks of 64 Ki values: byte counts, offsets, then the bytes */
  const int64_t chunk = 65536, nch = (len + chunk - 1) / chunk;
  int64_t* off = (int64_t*)calloc((size_t)nch + 1, sizeof(int64_t));
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int64_t c = 0; c < nch; ++c) {
    int64_t b = 0;
    const int64_t i1 = (c + 1) * chunk < len ? (c + 1) * chunk : len;
    for (int64_t i = c * chunk; i < i1; ++i) {
      if (i) ++b;
      put_sci(NULL, &b, f64 ? ((const double*)a)[i] : (double)((const float*)a)[i], digits);
    }
    off[c + 1] = b;
  }
  for (int64_t c = 0; c < nch; ++c) off[c + 1] += off[c];
  if (out) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int64_t c = 0; c < nch; ++c) {
      int64_t b = at + off[c];
      const int64_t i1 = (c + 1) * chunk < len ? (c + 1) * chunk : len;
      for (int64_t i = c * chunk; i < i1; ++i) {
        if (i) puts_(out, &b, ",");
        put_sci(out, &b, f64 ? ((const double*)a)[i] : (double)((const float*)a)[i], digits);
      }
    }
  }
  at += off[nch];
  free(off);
  return at;
}

static int64_t frame_pass(const int64_t n[3], const float* const vel[3], const float* tracer, const double* fraction,
                          const double* pressure, const or_d3_meta* m, char* out, int threads) {
  int64_t at = 0;
  puts_(out, &at, "{\"frame\":");
  put_u64(out, &at, m->frame);
  puts_(out, &at, ",\"time_s\":");
  put_sci(out, &at, m->time, 18);
  puts_(out, &at, ",\"dt_s\":");
  put_sci(out, &at, m->has_report ? m->dt : 0.0, 18);
  puts_(out, &at, ",\"carrier_stamp\":{\"id\":\"");
  put_u64(out, &at, m->carrier_id);
  puts_(out, &at, "\",\"version\":\"");
  put_u64(out, &at, m->carrier_version);
  puts_(out, &at, "\"},\"liquid_stamp\":{\"id\":\"");
  put_u64(out, &at, m->liquid_id);
  puts_(out, &at, "\",\"version\":\"");
  put_u64(out, &at, m->liquid_version);
  puts_(out, &at, "\"},\"fields\":{");
  const int64_t cells = n[0] * n[1] * n[2];
  static const char* const fnames[4] = {"velocity_x", "velocity_y", "velocity_z", "tracer"};
  for (int q = 0; q < 4; ++q) {
    int64_t len = cells;
    if (q < 3) len = (n[0] + (q == 0)) * (n[1] + (q == 1)) * (n[2] + (q == 2));
    puts_(out, &at, "\"");
    puts_(out, &at, fnames[q]);
    puts_(out, &at, "\":[");
    at = put_array(out, at, q < 3 ? (const void*)vel[q] : (const void*)tracer, 0, len, threads);
    puts_(out, &at, "],");
  }
  static const char* const dnames[2] = {"fraction", "pressure"};
  for (int q = 0; q < 2; ++q) {
    if (q) puts_(out, &at, ",");
    puts_(out, &at, "\"");
    puts_(out, &at, dnames[q]);
    puts_(out, &at, "\":[");
    at = put_array(out, at, q == 0 ? fraction : pressure, 1, cells, threads);
    puts_(out, &at, "]");
  }
  puts_(out, &at, "},\"diagnostics\":");
  if (m->has_report) {
    static const char* const keys[11] = {"pressure_residual_m3_s2", "carrier_divergence_s_inv", "liquid_divergence_s_inv",
                                         "volume_before_m3",        "volume_after_m3",          "mass_after_kg",
                                         "inward_m3",               "outward_m3",               "source_m3",
                                         "balance_m3",              "rounding_budget_m3"};
    puts_(out, &at, "{\"pressure_iterations\":");
    put_u64(out, &at, m->pressure_iterations);
    for (int k = 0; k < 11; ++k) {
      puts_(out, &at, ",\"");
      puts_(out, &at, keys[k]);
      puts_(out, &at, "\":");
      put_sci(out, &at, m->diag[k], 18);
    }
    puts_(out, &at, "}");
  } else {
    puts_(out, &at, "null");
  }
  puts_(out, &at, "}\n");
  return at;
}

int64_t or_dense3d_frame(const int64_t n[3], const float* u, const float* v, const float* w, const float* tracer,
                         const double* fraction, const double* pressure, const or_d3_meta* meta, char* out,
                         int64_t cap, int threads) {
  const float* const vel[3] = {u, v, w};
  const int64_t need = frame_pass(n, vel, tracer, fraction, pressure, meta, NULL, threads);
  if (!out || cap < need) return -need;
  return frame_pass(n, vel, tracer, fraction, pressure, meta, out, threads);
}
