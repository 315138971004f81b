// paging.hpp -- registers larger than the device budget.
//
// The state lives in pinned host memory as 2^K pieces of 2^nl amplitudes: the
// top K physical qubits index the piece, the low nl are the piece's own.
// A flush is cut into sections; in a section every non-diagonal target is a
// piece qubit, so one plan (controls and diagonals on the high qubits resolved
// per piece through dq_job.base_hi) runs on every piece as it streams
// host -> device -> host on three streams (copy in, compute, copy out).
// When the next gates target a high qubit, up to two high qubits trade places
// with the top piece qubits; the move is folded into the next section's copy
// in (each new piece gathered from the pieces of its group).  Pieces known to
// be zero are neither copied nor computed.
//
// Replaces the reference's single-GPU page executor
// (src/cpp/QuEST_backend_distributed.cpp:60-73 pinned pool, :123-199 task
// layout, :248-287 page tables, :289-647 prefetch / compute / drain with
// OPTIMIZE_ZERO) for one GPU without MPI.
#pragma once
#include <cstdint>
#include <vector>

#include "runtime.hpp"

namespace dgq {

struct PagedStats {
    int sections = 0;         // sections executed (each one pass over the pieces)
    int remaps = 0;           // high <-> piece qubit exchanges
    long long pieces_run = 0;     // pieces copied in, computed and copied out
    long long pieces_skipped = 0; // zero pieces left alone
    double copy_ms = 0;       // summed copy time (in + out)
    double hidden_ms = 0;     // copy time that overlapped compute
    double compute_ms = 0;    // summed kernel time
    double wall_ms = 0;       // flush wall time
    unsigned long long bytes_in = 0, bytes_out = 0;
};

struct PagedStore {
    int n = 0, nl = 0, K = 0, prec = DQ_F64;
    size_t elem = 16;
    size_t piece_bytes = 0;
    std::vector<unsigned char *> host;  // 2^K pinned pieces
    std::vector<uint8_t> zero;          // piece known to be all zeros
    std::vector<void *> dbuf;           // device piece buffers (two groups of up to 4)
    void *s_in = nullptr, *s_comp = nullptr, *s_out = nullptr;
    void *d_ws = nullptr;
    size_t ws_cap = 0;
    PagedStats st;
};

// Device bytes a register may use: DGQ_HBM_BUDGET (bytes) when set, else the
// device's free memory less a margin.
size_t device_budget();

// A paged store for an n-qubit register when its state exceeds `budget`
// bytes of device memory (nullptr otherwise).  The state starts as |0..0>.
PagedStore *paged_create(int n, int prec, size_t budget);
void paged_destroy(PagedStore *p);

void paged_flush(QuregState *s);   // run s->pending (and a pending init)
void paged_probs(QuregState *s);   // s->p0: P(q = 0) per physical qubit, total at [n]
// logical indices start .. start+count-1 as {re, im} doubles
void paged_read(QuregState *s, uint64_t start, uint64_t count, double *out);
// logical indices start .. start+count-1 from separate real / imaginary arrays
void paged_write(QuregState *s, uint64_t start, uint64_t count, const double *re, const double *im);

}  // namespace dgq
