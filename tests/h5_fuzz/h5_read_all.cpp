// Test harness (CPU only): opens every file named on the command line with the
// library's HDF5 reader (csrc/h5min.cpp, compiled into this program with
// AddressSanitizer by tests/test_h5_malformed.py) and reads every dataset path
// given after "--" as FP64, int64 and string.  Prints one status line per file;
// a read outside the file aborts under AddressSanitizer.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dfn.h"

int main(int argc, char **argv) {
  std::vector<const char *> files, paths;
  bool after = false;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--")) { after = true; continue; }
    (after ? paths : files).push_back(argv[i]);
  }
  std::vector<double> f64(1 << 16);
  std::vector<int64_t> i64(1 << 16);
  char str[256];
  for (const char *fn : files) {
    dfn_h5 h = nullptr;
    int st = dfn_h5_open(fn, &h);
    std::string line = std::to_string(st);
    if (st == DFN_OK) {
      for (const char *p : paths) {
        int32_t cls = 0, rank = 0;
        int64_t dims[8] = {0};
        int s = dfn_h5_info(h, p, &cls, &rank, dims);
        line += " " + std::to_string(s);
        if (s != DFN_OK) continue;
        int64_t n = 1;
        for (int r = 0; r < rank && r < 8; r++) n *= dims[r];
        if (n > 0 && n <= (int64_t)f64.size()) {
          line += "/" + std::to_string(dfn_h5_read_f64(h, p, f64.data(), n));
          line += "/" + std::to_string(dfn_h5_read_i64(h, p, i64.data(), n));
        }
        line += "/" + std::to_string(dfn_h5_read_string(h, p, str, (int32_t)sizeof(str)));
      }
      dfn_h5_close(h);
    }
    std::printf("%s\n", line.c_str());
  }
  return 0;
}
