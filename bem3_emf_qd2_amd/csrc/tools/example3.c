/* example3 -- instantaneous field images: the real part of E and H components times
 * exp(-i omega t) over one period (40 phases) on the planes x = 0, y = 0 and z = 0 (200 x 200
 * points over +-lambda0), written as RGB PNG files <dat name>_images/<plane>_<E|H><x|y|z>_NNN.png
 * with the colour bar, and <plane>_info.txt with each component's range (the output of the
 * reference's example3.c:41-439).  Each plane is one batched EH_t launch (bq_EH_t_batch).
 * The colour scale is the reference's nine-colour table, interpolated the same way
 * (example3.c:27-29, 403-439), so the pixels agree; the PNG files are written by the encoder
 * below (zlib stream of stored deflate blocks, no compression; libpng is not needed).
 * Optional arguments: points per side (default 200) and phases per period (default 40). */
#include <errno.h>
#include <stdint.h>
#include <sys/stat.h>
#include "bem3_emf_qd2.h"
#include "bq_amd.h"

/* ---- PNG writer (RFC 2083 / ISO 15948; zlib RFC 1950, deflate stored blocks RFC 1951) ---- */
static uint32_t crc_table[256];

static void crc_init(void)
{
  for (uint32_t n = 0; n < 256; n++) {
    uint32_t c = n;
    for (int k = 0; k < 8; k++) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
    crc_table[n] = c;
  }
}

static uint32_t crc_update(uint32_t c, const unsigned char *p, size_t n)
{
  for (size_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xff] ^ (c >> 8);
  return c;
}

static void put32(unsigned char *p, uint32_t v)
{
  p[0] = (unsigned char)(v >> 24);
  p[1] = (unsigned char)(v >> 16);
  p[2] = (unsigned char)(v >> 8);
  p[3] = (unsigned char)v;
}

static void chunk(FILE *fp, const char *type, const unsigned char *data, size_t n)
{
  unsigned char b[4];
  put32(b, (uint32_t)n);
  fwrite(b, 1, 4, fp);
  fwrite(type, 1, 4, fp);
  if (n) fwrite(data, 1, n, fp);
  uint32_t c = crc_update(0xffffffffu, (const unsigned char *)type, 4);
  c = crc_update(c, data, n) ^ 0xffffffffu;
  put32(b, c);
  fwrite(b, 1, 4, fp);
}

/* rgb: height rows of width * 3 bytes */
static void write_png(const char *fname, int width, int height, const unsigned char *rgb)
{
  FILE *fp = fopen(fname, "wb");
  if (!fp) {
    printf("Failed to open the %s file. Exit...\n", fname);
    exit(1);
  }
  static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
  fwrite(sig, 1, 8, fp);
  unsigned char ihdr[13];
  put32(ihdr, (uint32_t)width);
  put32(ihdr + 4, (uint32_t)height);
  ihdr[8] = 8;  /* bit depth */
  ihdr[9] = 2;  /* colour type RGB */
  ihdr[10] = 0; /* deflate */
  ihdr[11] = 0; /* filter method 0 */
  ihdr[12] = 0; /* no interlace */
  chunk(fp, "IHDR", ihdr, 13);
  /* raw scanlines: filter byte 0 + row */
  const size_t row = (size_t)width * 3 + 1, raw_n = row * (size_t)height;
  unsigned char *raw = malloc(raw_n);
  for (int y = 0; y < height; y++) {
    raw[y * row] = 0;
    memcpy(raw + y * row + 1, rgb + (size_t)y * width * 3, (size_t)width * 3);
  }
  /* zlib: header, stored blocks of at most 65535 bytes, Adler-32 */
  const size_t nblk = raw_n / 65535 + 1;
  unsigned char *z = malloc(2 + raw_n + 5 * nblk + 4);
  size_t o = 0;
  z[o++] = 0x78;
  z[o++] = 0x01;
  uint32_t a = 1, b = 0;
  for (size_t i = 0; i < raw_n || i == 0; i += 65535) {
    const size_t len = raw_n - i < 65535 ? raw_n - i : 65535;
    z[o++] = (i + len >= raw_n) ? 1 : 0;
    z[o++] = (unsigned char)(len & 0xff);
    z[o++] = (unsigned char)(len >> 8);
    z[o++] = (unsigned char)(~len & 0xff);
    z[o++] = (unsigned char)((~len >> 8) & 0xff);
    memcpy(z + o, raw + i, len);
    o += len;
    if (len == 0) break;
  }
  for (size_t i = 0; i < raw_n; i++) {
    a = (a + raw[i]) % 65521u;
    b = (b + a) % 65521u;
  }
  put32(z + o, (b << 16) | a);
  o += 4;
  chunk(fp, "IDAT", z, o);
  chunk(fp, "IEND", NULL, 0);
  fclose(fp);
  free(raw);
  free(z);
}

/* ---- colour scale -1 .. 1 (black outside) ---- */
static const unsigned char ct1[9][3] = {{0x00, 0x00, 0x90}, {0x00, 0x0f, 0xff}, {0x00, 0x90, 0xff},
                                        {0x0f, 0xff, 0xee}, {0xff, 0xff, 0xff}, {0xff, 0xee, 0x00},
                                        {0xff, 0x70, 0x00}, {0xee, 0x00, 0x00}, {0x7f, 0x00, 0x00}};

static void color_rgb(double x, unsigned char *rgb)
{
  if (!(x >= -1.0 && x <= 1.0)) {
    rgb[0] = rgb[1] = rgb[2] = 0;
    return;
  }
  /* 2^23 steps over [-1, 1], 8 segments of 2^21 steps between neighbouring table colours */
  const unsigned int n = (unsigned int)floor(8388608.0 * (x + 1.0)), nc = 2097152u;
  const unsigned int i = n / nc + 1;
  if (i > 8) {
    memcpy(rgb, ct1[8], 3);
    return;
  }
  const unsigned int nd = n - nc * (i - 1);
  for (int c = 0; c < 3; c++) {
    const double d = (double)(ct1[i][c] - ct1[i - 1][c]) * (1.0 / (double)nc);
    rgb[c] = (unsigned char)floor((double)ct1[i - 1][c] + d * (double)nd);
  }
}

typedef struct {
  char dir[300];
  int m, ts, scale;
  double rang;
  double complex *ve, *vh; /* [m][m][3] */
  double me[3], mh[3];
} Img;

/* plane 0: x = 0 (rows z from +rang down, columns y); 1: y = 0 (rows z, columns x);
   2: z = 0 (rows y from +rang down, columns x) */
static void fields(Img *id, DQD2 *qd, int plane)
{
  const int m = id->m;
  const double dr = id->rang * 2.0 / (double)(m - 1);
  const size_t n = (size_t)m * m;
  double *r = calloc(n * 3, sizeof(double));
  int *did = malloc(n * sizeof(int));
  const int fixed = plane, row = plane == 2 ? 1 : 2, col = plane == 0 ? 1 : 0;
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++) {
      double *p = r + 3 * ((size_t)i * m + j);
      p[fixed] = 0.0;
      p[row] = id->rang - (double)i * dr;
      p[col] = -id->rang + (double)j * dr;
    }
  if (bq_EH_t_batch(qd, (int)n, r, 1, (double *)id->ve, (double *)id->vh, did) != 0) {
    printf("example3: field evaluation failed: %s\n", bq_last_error());
    exit(1);
  }
  for (int d = 0; d < 3; d++) id->me[d] = id->mh[d] = 0.0;
  for (size_t k = 0; k < n; k++)
    for (int d = 0; d < 3; d++) {
      if (cabs(id->ve[3 * k + d]) > id->me[d]) id->me[d] = cabs(id->ve[3 * k + d]);
      if (cabs(id->vh[3 * k + d]) > id->mh[d]) id->mh[d] = cabs(id->vh[3 * k + d]);
    }
  free(r);
  free(did);
}

static void output_field(const char *pl, Img *id, DQD2 *qd)
{
  const int m = id->m, s1 = id->scale + 1, w = m * s1;
  const double dt = qd->pw.lambda_0 / (double)id->ts;
  const char sf[3] = {'x', 'y', 'z'};
  unsigned char *img = malloc((size_t)w * w * 3);
  char fname[512];
  for (int nt = 0; nt < id->ts; nt++) {
    const double complex cet = cexp(-I * qd->pw.omega * dt * (double)nt);
    for (int f = 0; f < 2; f++)
      for (int d = 0; d < 3; d++) {
        const double complex *v = f == 0 ? id->ve : id->vh;
        const double mx = f == 0 ? id->me[d] : id->mh[d];
        for (int i = 0; i < m; i++)
          for (int j = 0; j < m; j++) {
            unsigned char c[3];
            color_rgb(creal(cet * v[(size_t)i * m * 3 + j * 3 + d]) / mx, c);
            for (int si = 0; si < s1; si++)
              for (int sj = 0; sj < s1; sj++) memcpy(img + ((size_t)(i * s1 + si) * w + j * s1 + sj) * 3, c, 3);
          }
        snprintf(fname, sizeof fname, "%s/%s_%c%c_%03d.png", id->dir, pl, f == 0 ? 'E' : 'H', sf[d], nt);
        write_png(fname, w, w, img);
      }
  }
  free(img);
  snprintf(fname, sizeof fname, "%s/%s_info.txt", id->dir, pl);
  FILE *fp = fopen(fname, "wt");
  if (!fp) {
    printf("Failed to open the %s file. Exit...\n", fname);
    exit(1);
  }
  fprintf(fp, "the range of color bar\n");
  fprintf(fp, "Ex is %8e to %8e\n", -id->me[0], id->me[0]);
  fprintf(fp, "Ey is %8e to %8e\n", -id->me[1], id->me[1]);
  fprintf(fp, "Ez is %8e to %8e\n", -id->me[2], id->me[2]);
  fprintf(fp, "Hx is %8e to %8e\n", -id->mh[0], id->mh[0]);
  fprintf(fp, "Hy is %8e to %8e\n", -id->mh[1], id->mh[1]);
  fprintf(fp, "Hz is %8e to %8e\n", -id->mh[2], id->mh[2]);
  fclose(fp);
  /* colour bar: +1 at the top to -1 at the bottom */
  const int h = w, bw = h / 16 > 0 ? h / 16 : 1;
  unsigned char *bar = malloc((size_t)bw * h * 3);
  for (int i = 0; i < h; i++) {
    unsigned char c[3];
    color_rgb(1.0 - (2.0 / (double)h) * (double)i, c);
    for (int j = 0; j < bw; j++) memcpy(bar + ((size_t)i * bw + j) * 3, c, 3);
  }
  snprintf(fname, sizeof fname, "%s/color_bar.png", id->dir);
  write_png(fname, bw, h, bar);
  free(bar);
}

int main(int argc, char **argv)
{
  DQD2 qd;
  Img id;
  if (argc < 2) {
    printf("usage: %s ex.dat [points_per_side [phases]]\n", argv[0]);
    return 1;
  }
  id.m = argc > 2 ? atoi(argv[2]) : 200;
  id.ts = argc > 3 ? atoi(argv[3]) : 40;
  id.scale = 1;
  if (id.m < 2 || id.ts < 1) {
    printf("example3: points per side must be at least 2 and phases at least 1\n");
    return 1;
  }
  crc_init();
  dat_read(argv[1], &qd);
  print_dqd2(&qd);
  /* <name without extension>_images */
  snprintf(id.dir, sizeof id.dir, "%s", argv[1]);
  char *dot = strrchr(id.dir, '.'), *slash = strrchr(id.dir, '/');
  if (dot && (!slash || dot > slash)) *dot = '\0';
  strncat(id.dir, "_images", sizeof id.dir - strlen(id.dir) - 1);
  if (mkdir(id.dir, S_IRWXU | S_IRWXG) != 0 && errno != EEXIST) {
    printf("failed to make directory. Exit..");
    exit(1);
  }
  id.rang = 1.0 * qd.pw.lambda_0;
  id.ve = malloc((size_t)id.m * id.m * 3 * sizeof(double complex));
  id.vh = malloc((size_t)id.m * id.m * 3 * sizeof(double complex));
  const char *pl[3] = {"yz", "xz", "xy"};
  for (int p = 0; p < 3; p++) {
    fields(&id, &qd, p);
    output_field(pl[p], &id, &qd);
  }
  free(id.ve);
  free(id.vh);
  finalize_dqd2(&qd);
  return 0;
}
