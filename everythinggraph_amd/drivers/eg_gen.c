/*
 * eg_gen -- writes a synthetic binary edge file ([src][dst] u32 records, or
 * [src][dst][f32 w] with -w), the input format of init_all.h:8-14.  The
 * reference has no generator; records come from libeg's device generator.
 *
 *   eg_gen -k rmat -s SCALE -e EDGE_FACTOR [-S seed] [-w] [-W wseed] -o FILE
 *   eg_gen -k er -n N -m EDGES [-S seed] [-w] [-W wseed] -o FILE
 * R-MAT uses a,b,c = 0.57, 0.19, 0.19 unless -a/-b/-c are given.
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <stdio.h>

#include "eg.h"

static void eg_check(int rc, const char *what) {
  if (rc != EG_OK) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, eg_last_error());
    exit(1);
  }
}

static void gen_usage(void) {
  fprintf(stderr,
          "usage: eg_gen -k rmat -s SCALE -e EF | -k er -n N -m EDGES  [-S seed] [-w] [-W wseed] "
          "[-a A -b B -c C] -o FILE\n");
  exit(2);
}

int main(int argc, char **argv) {
  eg_gen_params p;
  memset(&p, 0, sizeof(p));
  p.a = 0.57;
  p.b = 0.19;
  p.c = 0.19;
  p.seed = 1;
  p.weight_seed = 1;
  const char *out = NULL;
  int scale = 0, ef = 16, weighted = 0;
  uint32_t n = 0;
  uint64_t m = 0;
  int c;
  while ((c = getopt(argc, argv, "k:s:e:n:m:S:wW:a:b:c:o:")) != -1) {
    switch (c) {
      case 'k': p.kind = strcmp(optarg, "er") == 0 ? 1 : 0; break;
      case 's': scale = atoi(optarg); break;
      case 'e': ef = atoi(optarg); break;
      case 'n': n = (uint32_t)strtoul(optarg, NULL, 10); break;
      case 'm': m = strtoull(optarg, NULL, 10); break;
      case 'S': p.seed = strtoull(optarg, NULL, 10); break;
      case 'w': weighted = 1; break;
      case 'W': p.weight_seed = strtoull(optarg, NULL, 10); break;
      case 'a': p.a = atof(optarg); break;
      case 'b': p.b = atof(optarg); break;
      case 'c': p.c = atof(optarg); break;
      case 'o': out = optarg; break;
      default: gen_usage();
    }
  }
  if (!out) gen_usage();
  if (p.kind == 0) {
    if (scale < 1 || scale > 31) gen_usage();
    p.scale = scale;
    n = 1u << scale;
    m = (uint64_t)n * (uint64_t)ef;
  } else if (!n || !m) {
    gen_usage();
  }
  p.m_in = m;
  uint32_t *src = (uint32_t *)malloc(m * 4), *dst = (uint32_t *)malloc(m * 4);
  float *w = weighted ? (float *)malloc(m * 4) : NULL;
  if (!src || !dst || (weighted && !w)) {
    fprintf(stderr, "out of host memory\n");
    return 1;
  }
  eg_check(eg_generate_edges(&p, n, src, dst, w), "generate");
  FILE *f = fopen(out, "wb");
  if (!f) {
    perror(out);
    return 1;
  }
  const size_t chunk = 1 << 20;
  uint32_t *buf = (uint32_t *)malloc(chunk * 3 * 4);
  for (uint64_t i = 0; i < m; i += chunk) {
    const size_t cnt = m - i < chunk ? (size_t)(m - i) : chunk;
    size_t k = 0;
    for (size_t j = 0; j < cnt; j++) {
      buf[k++] = src[i + j];
      buf[k++] = dst[i + j];
      if (w) memcpy(&buf[k++], &w[i + j], 4);
    }
    if (fwrite(buf, 4, k, f) != k) {
      perror("write");
      return 1;
    }
  }
  fclose(f);
  printf("wrote %llu records (n = %u) to %s\n", (unsigned long long)m, n, out);
  free(buf);
  free(src);
  free(dst);
  free(w);
  return 0;
}
