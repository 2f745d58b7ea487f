// Test harness: routes correlator-node slices (integration_write bytes, one
// record per bin or phase centre) into per-file .cor files through
// Output_node_writer (sfxc-1_amd/host/sfxc_output.h), in pieces of a given
// length. Built by tests/test_corfile.py with g++.
//
// Usage: cor_route <prefix> <piece> <n_records> <nchan> <slice file>...
// writes <prefix>.<file index>.cor for every file index the records carry.
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "../../sfxc-1_amd/host/sfxc_output.h"

class File_writer : public Data_writer {
 public:
  explicit File_writer(const std::string &p) : f_(fopen(p.c_str(), "wb")) {
    if (!f_) throw std::runtime_error("cannot open " + p);
  }
  ~File_writer() override { fclose(f_); }

 protected:
  size_t do_put_bytes(size_t n, const char *b) override { return fwrite(b, 1, n, f_); }

 private:
  FILE *f_;
};

static std::string g_prefix;
static std::shared_ptr<Data_writer> open_file(int32_t index, void *) {
  return std::make_shared<File_writer>(g_prefix + "." + std::to_string(index) + ".cor");
}

int main(int argc, char **argv) {
  if (argc < 6) return 2;
  g_prefix = argv[1];
  const long piece = atol(argv[2]);
  const int n_records = atoi(argv[3]);
  Output_header_global h;
  memset(&h, 0, sizeof(h));
  strncpy(h.experiment, "N26X1", sizeof(h.experiment));
  h.start_year = 2023;
  h.start_day = 55;
  h.number_channels = atoi(argv[4]);
  h.integration_time = 1000000;
  h.output_format_version = 1;
  h.polarisation_type = 3;
  try {
    Output_node_writer out(h, open_file, nullptr);
    for (int a = 5; a < argc; a++) {
      FILE *f = fopen(argv[a], "rb");
      if (!f) return 3;
      std::vector<char> buf;
      char tmp[65536];
      size_t k;
      while ((k = fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + k);
      fclose(f);
      out.set_slice((int64_t)buf.size(), n_records);
      for (size_t o = 0; o < buf.size(); o += (size_t)piece)
        if (!out.write_output(buf.data() + o, (int64_t)std::min(buf.size() - o, (size_t)piece))) return 4;
    }
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
