// Bit writing and NAL packaging shared by the slice writers (host code): an MSB-first bit
// writer with Exp-Golomb codes (7.2), NAL units with start codes and emulation prevention
// (enc.rs:2643-2672), and the copy into a caller's buffer.
#pragma once
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/ragdoll_hip.h"

namespace rgh_bits {

class BitWriter {
  public:
    void put(uint32_t value, int len) {  // len <= 32
        for (int i = len - 1; i >= 0; i--) bit((value >> i) & 1u);
    }
    void zeros(int n) {
        for (int i = 0; i < n; i++) bit(0);
    }
    void ue(uint32_t v) {
        const uint64_t x = (uint64_t)v + 1;
        int len = 0;
        while ((x >> len) > 1) len++;
        zeros(len);
        for (int i = len; i >= 0; i--) bit((uint32_t)(x >> i) & 1u);
    }
    void se(int v) { ue(v > 0 ? 2u * (uint32_t)v - 1u : 2u * (uint32_t)(-v)); }
    void trailing() {  // rbsp_stop_one_bit and alignment
        bit(1);
        while (fill_ != 0) bit(0);
    }
    void align_ones() {  // cabac_alignment_one_bit
        while (fill_ != 0) bit(1);
    }
    bool aligned() const { return fill_ == 0; }
    std::vector<uint8_t> &bytes() { return buf_; }
    const std::vector<uint8_t> &bytes() const { return buf_; }

  private:
    void bit(uint32_t b) {
        cur_ = (uint8_t)((cur_ << 1) | b);
        if (++fill_ == 8) {
            buf_.push_back(cur_);
            cur_ = 0;
            fill_ = 0;
        }
    }
    std::vector<uint8_t> buf_;
    uint8_t cur_ = 0;
    int fill_ = 0;
};

// NAL unit (enc.rs:2643): start code, header byte, emulation prevention; appended to nal.
inline void emit_nal(std::vector<uint8_t> &nal, int ref_idc, int type, const std::vector<uint8_t> &rbsp) {
    nal.insert(nal.end(), {0, 0, 0, 1, (uint8_t)((ref_idc << 5) | type)});
    int zeros = 0;
    for (uint8_t b : rbsp) {
        if (zeros == 2 && b <= 3) {
            nal.push_back(3);
            zeros = 0;
        }
        nal.push_back(b);
        zeros = b == 0 ? zeros + 1 : 0;
    }
}

inline rgh_status deliver(const std::vector<uint8_t> &nal, uint8_t *out, int64_t cap, int64_t *out_len) {
    *out_len = (int64_t)nal.size();
    if ((int64_t)nal.size() > cap) return RGH_ERR_INVALID_ARGUMENT;
    memcpy(out, nal.data(), nal.size());
    return RGH_OK;
}

}  // namespace rgh_bits
