// internal.h -- declarations shared by the library's translation units.
#ifndef SNAPMRF_INTERNAL_H
#define SNAPMRF_INTERNAL_H

#include <cstddef>
#include <cstdint>

// Sets the thread-local error message and returns `code`.
int mrf_fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void mrf_clear_error(void);

namespace mrf {

enum Echo { kEpgSsfp = 0, kRoaSsfp = 1, kEpgSr = 2, kEpgSe = 3, kEpgTse = 4, kEpgFisp = 5 };

// Echo-type string -> program id (src/functions.cu:1238-1311); -1 if unknown.
int echo_id(const char *echo_type);

// Per-TR coefficient record consumed by the EPG/ROA kernels (one per TR per
// B1 slice).  All fields are float32 values derived on the host in double
// from the prepared float32 sequence.
struct TrCoef {
    float ca2, sa2, sa, ca;   // (1+cos a)/2, (1-cos a)/2, sin a, cos a of alpha*B1
    float zr, zi;             // e^{i phi_n}
    float rr, ri;             // e^{i (phi_n - phi_{n+1})} (phi_T := phi_{T-1})
    float te, tr, tmte;       // TE_n, TR_n, TR_n - TE_n (float32, as the reference)
    float s2p, c2p;           // sin(2 phi_n), cos(2 phi_n) (ROA rotation)
    float pad[3];             // pad[0]: alpha * B1 itself
};
static_assert(sizeof(TrCoef) == 64, "TrCoef must be 64 bytes");

}  // namespace mrf

#endif
