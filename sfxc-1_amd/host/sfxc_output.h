// sfxc_output.h -- the output node's end of the correlator-node data path,
// for hosts built outside the SFXC tree: the global header of a .cor file
// and the routing of a slice's records to per-bin (or per-source) files.
//
//   Output_header_global   include/output_header.h:70-96 (84 bytes, laid out
//                          as the reference's struct on x86-64)
//   Output_node_writer     Output_node::write_output (src/output_node.cc:
//                          181-213): a slice holds n_bins records of equal
//                          length, each opened by an int32 file index
//                          (integration_write's bin or source number); the
//                          index is taken off and the rest of the record is
//                          appended to that file. Bytes may arrive in pieces
//                          of any length.
//
// The global header's contents (experiment, dates, versions, job numbers)
// come from the manager node (src/manager_node.cc:670-703); the caller fills
// them in.
#pragma once

#include <stdint.h>
#include <string.h>

#include <map>
#include <memory>
#include <stdexcept>

#include "sfxc_types.h"

#pragma pack(push, 1)
struct Output_header_global {
  int32_t header_size;
  char experiment[32];
  int16_t start_year;
  int16_t start_day;
  int32_t start_time;        // seconds since midnight of the start day
  int32_t number_channels;
  int32_t integration_time;  // microseconds
  int32_t output_format_version;
  int32_t correlator_version;
  int8_t polarisation_type;  // 0 L, 1 R, 2 L+R, 3 L+R with crosses
  char correlator_branch[15];
  int32_t job_nr;
  int32_t subjob_nr;
};
#pragma pack(pop)
static_assert(sizeof(Output_header_global) == 84, "Output_header_global layout");

class Output_node_writer {
 public:
  // Opens the writer of a file index on first use; the global header goes
  // first into every file.
  typedef std::shared_ptr<Data_writer> (*Writer_factory)(int32_t file_index, void *user);

  Output_node_writer(const Output_header_global &header, Writer_factory factory, void *user)
      : header_(header), factory_(factory), user_(user) {
    header_.header_size = (int32_t)sizeof(Output_header_global);
  }

  // The next slice holds n_records records in slice_size bytes.
  void set_slice(int64_t slice_size, int32_t n_records) {
    if (n_records <= 0 || slice_size % n_records != 0 || slice_size / n_records <= 4)
      throw std::invalid_argument("a slice holds whole records of more than their 4-byte index");
    if (record_left_ != 0 || index_have_ != 0) throw std::logic_error("the previous slice is incomplete");
    record_len_ = slice_size / n_records;
    slice_left_ = slice_size;
  }

  // n bytes of the slice, in order; returns false if they overrun it.
  bool write_output(const char *buf, int64_t n) {
    while (n > 0) {
      if (slice_left_ <= 0) return false;
      if (record_left_ == 0) {  // the record's file index, possibly split over pieces
        const int64_t k = std::min<int64_t>(4 - index_have_, n);
        memcpy(index_bytes_ + index_have_, buf, (size_t)k);
        index_have_ += (int)k;
        buf += k;
        n -= k;
        slice_left_ -= k;
        if (index_have_ < 4) continue;
        memcpy(&file_, index_bytes_, 4);
        index_have_ = 0;
        record_left_ = record_len_ - 4;
        continue;
      }
      const int64_t k = std::min(record_left_, n);
      writer(file_)->put_bytes((size_t)k, buf);
      buf += k;
      n -= k;
      record_left_ -= k;
      slice_left_ -= k;
    }
    return true;
  }

 private:
  Data_writer *writer(int32_t file) {
    auto it = files_.find(file);
    if (it == files_.end()) {
      std::shared_ptr<Data_writer> w = factory_(file, user_);
      if (!w) throw std::runtime_error("no writer for output file " + std::to_string(file));
      w->put_bytes(sizeof(header_), reinterpret_cast<const char *>(&header_));
      it = files_.emplace(file, w).first;
    }
    return it->second.get();
  }

  Output_header_global header_;
  Writer_factory factory_;
  void *user_;
  std::map<int32_t, std::shared_ptr<Data_writer>> files_;
  int64_t record_len_ = 0, record_left_ = 0, slice_left_ = 0;
  char index_bytes_[4] = {0, 0, 0, 0};
  int index_have_ = 0;
  int32_t file_ = 0;
};
