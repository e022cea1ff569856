// Host-only driver for the COG reader's IFD parser (tiff_ifd.hpp), built with
// the address and undefined-behaviour sanitizers by tests/test_tiff_parse.py:
// one line per file argument, "OK n_levels w h tw th" or "ERR reason".
#include <fcntl.h>
#include <unistd.h>

#include <cstdio>

#include "tiff_ifd.hpp"

int main(int argc, char** argv) {
  for (int i = 1; i < argc; i++) {
    const int fd = open(argv[i], O_RDONLY);
    if (fd < 0) {
      std::printf("ERR cannot open\n");
      continue;
    }
    fsg::TiffFile F;
    const std::string why = fsg::tiff_parse(fd, F);
    close(fd);
    if (why.empty())
      std::printf("OK %zu %d %d %d %d\n", F.lv.size(), F.lv[0].w, F.lv[0].h, F.lv[0].tw,
                  F.lv[0].th);
    else
      std::printf("ERR %s\n", why.c_str());
  }
  return 0;
}
