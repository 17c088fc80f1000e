/*
 * CPU oracle: target segments of a 2-bit .pac reference, restated from
 * FerrousAlign.  TEST INFRASTRUCTURE ONLY: only tests/, the smoke check and
 * bench.py's CPU baseline may call it; the library never does.
 *
 *   get_pac            src/pipelines/linear/index/bntseq.rs:559-565
 *   extract_forward    bntseq.rs:627-651 (extract_forward_batch; the byte-at-a-
 *                      time middle section gives the same bases as get_pac)
 *   get_segment        bntseq.rs:451-522 (get_reference_segment)
 *   left-job reversal  src/pipelines/linear/batch_extension/collect_soa.rs:156-176
 */
#include <stdint.h>

#include "ksw_oracle.h"

/* base p of the forward strand: byte p>>2, two bits at ((~p)&3)*2 */
static uint8_t get_pac(const uint8_t *pac, int64_t pos) {
    const int64_t byte_idx = pos >> 2;
    const unsigned shift = (unsigned)((~(uint8_t)pos) & 3u) << 1;
    return (uint8_t)((pac[byte_idx] >> shift) & 3u);
}

/* forward bases [beg, end): unaligned prefix, whole bytes (4 bases, first
 * base in the top bits), unaligned suffix */
static int64_t extract_forward(const uint8_t *pac, int64_t beg, int64_t end, uint8_t *out) {
    int64_t n = 0, pos = beg;
    if (beg >= end) return 0;
    while (pos < end && (pos & 3) != 0) out[n++] = get_pac(pac, pos++);
    const int64_t aligned_end = end & ~(int64_t)3;
    while (pos < aligned_end) {
        const uint8_t b = pac[pos >> 2];
        out[n++] = (b >> 6) & 3u;
        out[n++] = (b >> 4) & 3u;
        out[n++] = (b >> 2) & 3u;
        out[n++] = b & 3u;
        pos += 4;
    }
    while (pos < end) out[n++] = get_pac(pac, pos++);
    return n;
}

int64_t ora_pac_segment(const uint8_t *pac, int64_t l_pac, int64_t start, int64_t len,
                        int reverse, uint8_t *out) {
    const int64_t end = start + len;
    int64_t n;
    if (len < 0 || start < 0 || end > 2 * l_pac) return -1;
    if (start < l_pac && end > l_pac) return -1; /* bridging: the reference returns empty */
    if (start >= l_pac) {
        /* reverse-complement strand: forward [2 l_pac - end, 2 l_pac - start),
         * then reversed and complemented */
        const int64_t beg_f = 2 * l_pac - end, end_f = 2 * l_pac - start;
        n = extract_forward(pac, beg_f, end_f, out);
        for (int64_t i = 0, j = n - 1; i < j; ++i, --j) {
            const uint8_t t = out[i];
            out[i] = out[j];
            out[j] = t;
        }
        for (int64_t i = 0; i < n; ++i) out[i] = (uint8_t)(3 - out[i]);
    } else {
        n = extract_forward(pac, start, end, out);
    }
    if (reverse) /* left extension: the target segment read backwards */
        for (int64_t i = 0, j = n - 1; i < j; ++i, --j) {
            const uint8_t t = out[i];
            out[i] = out[j];
            out[j] = t;
        }
    return n;
}
