/*
 * Text I/O of the BdG API (reference src/io.c:13-165): "re im" wavefunction
 * loader, tab-separated %.8e eigenvalue rows, u/v mode writers with blank
 * lines between y-rows (2D) or z-slices (3D).
 */
#include <stdarg.h>

#include "bdg_host.h"

int bdg_load_wavefunction(bdg_t *b, const char *filename) {
  BDG_REQUIRE(b, BDG_HAS_SYSTEM, "bdg_load_wavefunction");
  const uint64_t S = b->size;
  FILE *f = fopen(filename, "r");
  if (!f) return -1;
  const int c = b->complex_psi0;
  double *wf = (double *)calloc(S * (c ? 2 : 1), sizeof(double));
  if (!wf) {
    fclose(f);
    return -1;
  }
  for (uint64_t i = 0; i < S; ++i) {
    double re, im;
    if (2 != fscanf(f, "%lf %lf", &re, &im)) {
      free(wf);
      fclose(f);
      return -2;
    }
    if (c) {
      wf[2 * i] = re;
      wf[2 * i + 1] = im;
    } else {
      wf[i] = re;
    }
  }
  fclose(f);
  bdg_set_wavefunction(b, wf, S);
  free(wf);
  return 0;
}

int bdg_load_wavefunction_fmt(bdg_t *b, const char *fmt, ...) {
  char name[512];
  va_list ap;
  va_start(ap, fmt);
  const int n = vsnprintf(name, sizeof(name), fmt, ap);
  va_end(ap);
  if (n < 0 || (size_t)n >= sizeof(name)) return -1;
  return bdg_load_wavefunction(b, name);
}

int bdg_write_eigenvalues(const bdg_t *b, const char *filename) {
  if (!b->eigvals) return -3;
  FILE *f = fopen(filename, "a");
  if (!f) return -1;
  for (uint64_t i = 0; i < b->nev; ++i) {
    fprintf(f, "%.8e", b->eigvals[i]);
    if (i + 1 < b->nev) fputc('\t', f);
  }
  fputc('\n', f);
  fclose(f);
  return 0;
}

static int write_mode(const bdg_t *b, uint64_t idx, const void *modes, const char *base,
                      const char *part) {
  if (!modes) return -3;
  if (idx >= b->nev) return -2;
  char name[512];
  snprintf(name, sizeof(name), "%s_%s_%lu.dat", base, part, (unsigned long)idx);
  FILE *f = fopen(name, "w");
  if (!f) return -1;
  const uint64_t S = b->size;
  const int c = b->complex_psi0;
  const double *md = (const double *)modes + idx * S * (c ? 2 : 1);
  const uint64_t nx = b->N[0], ny = b->dim > 1 ? b->N[1] : 1, nz = b->dim > 2 ? b->N[2] : 1;
  for (uint64_t k = 0; k < nz; ++k) {
    for (uint64_t j = 0; j < ny; ++j) {
      for (uint64_t i = 0; i < nx; ++i) {
        const uint64_t id = (k * ny + j) * nx + i;
        const double re = c ? md[2 * id] : md[id];
        const double im = c ? md[2 * id + 1] : 0.0;
        if (b->dim == 1) {
          fprintf(f, "%.8e\t%.8e\n", re, im);
        } else {
          fprintf(f, "%.8e\t%.8e", re, im);
          if (i + 1 < nx) fputc('\t', f);
        }
      }
      if (b->dim > 1) fputc('\n', f);
      if (b->dim == 2 && j + 1 < ny) fputc('\n', f);
    }
    if (b->dim == 3 && k + 1 < nz) fputc('\n', f);
  }
  fclose(f);
  return 0;
}

int bdg_write_mode_u(const bdg_t *b, uint64_t idx, const char *base) {
  return write_mode(b, idx, b->modes_u, base, "u");
}

int bdg_write_mode_v(const bdg_t *b, uint64_t idx, const char *base) {
  return write_mode(b, idx, b->modes_v, base, "v");
}

/* ---- binary fast path (no reference counterpart) ---------------------------
 * The text formats above write ~30 bytes per value through stdio; at 128^3
 * and beyond a wavefunction or a set of modes is hundreds of MB of text.
 * The binary files hold the same values bit for bit: a 64-byte header
 *   char magic[8] "BDGBIN1"; uint32 kind (1 wavefunction, 2 modes);
 *   uint32 complex; uint64 dim; uint64 N[3]; uint64 count (1, or nev);
 *   uint64 reserved
 * followed by little-endian FP64 payload (interleaved re, im when complex):
 * kind 1: psi0 (S values); kind 2: the nev eigenvalues (real), then modes_u
 * (S x nev, column-major), then modes_v. */
typedef struct {
  char magic[8];
  uint32_t kind, cplx;
  uint64_t dim, N[3], count, reserved;
} bdg_bin_header;

static const char kBinMagic[8] = "BDGBIN1";

static void bin_header(const bdg_t *b, uint32_t kind, uint64_t count, bdg_bin_header *h) {
  memset(h, 0, sizeof(*h));
  memcpy(h->magic, kBinMagic, sizeof(kBinMagic));
  h->kind = kind;
  h->cplx = (uint32_t)b->complex_psi0;
  h->dim = b->dim;
  for (int d = 0; d < 3; ++d) h->N[d] = d < (int)b->dim ? b->N[d] : 1;
  h->count = count;
}

static int write_all(FILE *f, const void *p, size_t bytes) {
  return fwrite(p, 1, bytes, f) == bytes ? 0 : -1;
}

int bdg_save_wavefunction_bin(const bdg_t *b, const char *filename) {
  if (!b->wf) return -3;
  FILE *f = fopen(filename, "wb");
  if (!f) return -1;
  bdg_bin_header h;
  bin_header(b, 1, 1, &h);
  const size_t bytes = b->size * sizeof(double) * (b->complex_psi0 ? 2 : 1);
  int rc = write_all(f, &h, sizeof(h));
  if (!rc) rc = write_all(f, b->wf, bytes);
  if (fclose(f) != 0) rc = -1;
  return rc;
}

int bdg_load_wavefunction_bin(bdg_t *b, const char *filename) {
  BDG_REQUIRE(b, BDG_HAS_SYSTEM, "bdg_load_wavefunction_bin");
  FILE *f = fopen(filename, "rb");
  if (!f) return -1;
  bdg_bin_header h, want;
  bin_header(b, 1, 1, &want);
  if (fread(&h, 1, sizeof(h), f) != sizeof(h) || memcmp(h.magic, kBinMagic, 8) != 0 ||
      h.kind != 1 || h.cplx != want.cplx || h.dim != want.dim || h.count != 1 ||
      memcmp(h.N, want.N, sizeof(h.N)) != 0) {
    fclose(f);
    return -2;
  }
  const size_t n = b->size * (b->complex_psi0 ? 2 : 1);
  double *wf = (double *)malloc(n * sizeof(double));
  if (!wf) {
    fclose(f);
    return -1;
  }
  const size_t got = fread(wf, sizeof(double), n, f);
  fclose(f);
  if (got != n) {
    free(wf);
    return -2;
  }
  bdg_set_wavefunction(b, wf, b->size);
  free(wf);
  return 0;
}

int bdg_write_modes_bin(const bdg_t *b, const char *filename) {
  if (!b->eigvals || !b->modes_u || !b->modes_v) return -3;
  FILE *f = fopen(filename, "wb");
  if (!f) return -1;
  bdg_bin_header h;
  bin_header(b, 2, b->nev, &h);
  const size_t mb = b->size * b->nev * sizeof(double) * (b->complex_psi0 ? 2 : 1);
  int rc = write_all(f, &h, sizeof(h));
  if (!rc) rc = write_all(f, b->eigvals, b->nev * sizeof(double));
  if (!rc) rc = write_all(f, b->modes_u, mb);
  if (!rc) rc = write_all(f, b->modes_v, mb);
  if (fclose(f) != 0) rc = -1;
  return rc;
}
