// Host build of the shared decimal <-> binary64 conversions (number.hpp,
// format.hpp).  Default: reads one token per line from stdin, prints the
// result bits in hex, or "E<rc>".  With argument "fmt": reads "style bits"
// lines (bits in hex), prints the formatted field between '|'.
#include <cstdio>
#include <cstring>
#include <string>

#include "../critic2_amd/csrc/format.hpp"

int main(int argc, char **argv) {
  using namespace c2;
  static uint64_t p5[2 * (POW5_QMAX - POW5_QMIN + 1)];
  static int16_t p5s[POW5_QMAX - POW5_QMIN + 1];
  pow5_table(p5, p5s);
  char line[8192];
  if (argc > 1 && !strcmp(argv[1], "fmt")) {
    int style;
    unsigned long long bits;
    while (scanf("%d %llx", &style, &bits) == 2) {
      double x;
      memcpy(&x, &bits, 8);
      char f[32] = {0};
      fmt_value(x, style, p5, p5s, f);
      f[fmt_width(style)] = 0;
      printf("|%s|\n", f);
    }
    return 0;
  }
  while (fgets(line, sizeof line, stdin)) {
    int n = (int)strlen(line);
    while (n > 0 && (line[n - 1] == '\n' || line[n - 1] == '\r')) n--;
    DecNum d;
    int rc = dec_parse(line, n, d);
    double v = 0;
    if (rc == 0) rc = dec_to_double(line, d, p5, p5s, v);
    if (rc) {
      printf("E%d\n", rc);
      continue;
    }
    uint64_t b;
    memcpy(&b, &v, 8);
    printf("%016llx\n", (unsigned long long)b);
  }
  return 0;
}
