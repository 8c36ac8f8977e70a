// Batched KEM chains (key generation, encapsulation, decapsulation) on the GPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace tmvp {

// Parameter set with the derived byte sizes (S/include/params.h:26-49,
// SB/include/SABLE_params.h:26-72).
struct KemParams {
    int l, eq, ep, et, mu, sable, h1, h2;
    int pvc;       // L * EP * 32: compressed polynomial vector
    int indcpa_sk; // Saber: L * 13 * 32, Sable: L * 2 * 32
    int pk, sk, ct;
    int araw; // L * L * EQ * 32: bytes of the SHAKE-128 stream GenMatrix unpacks
    int sraw; // L * MU * 32: bytes of the SHAKE-128 stream GenSecret samples from
};

// Scratch for one batched call (carved from one device allocation).
struct KemWorkspace {
    uint16_t *A, *s, *v1, *v2, *bits, *poly;
    uint8_t *buf, *kr;
    uint32_t *diff; // decapsulation: re-encryption differs from ct
    uint8_t *hct;              // H(ct) per instance (decapsulation)
    // Decapsulation runs the steps that depend only on its inputs (the matrix
    // stream, H(ct)) on `aux`, forked from and joined to the caller's stream
    // with the two events.  Owned by whoever owns the workspace.
    hipStream_t aux;
    hipEvent_t fork, join;
};

// Device-resident keys of one parameter set: pk [n][pk], sk [n][sk] (may be
// null), the matrices A [n][L][L][256] and H(pk) [n][32] derived from them.
struct KeyStore {
    const uint8_t *pk, *sk;
    const uint16_t *A;
    const uint8_t *hpk;
};

// Bytes of a workspace for B instances; with_A = false leaves out the matrix
// (W.A = null), which only keyed calls on the MFMA backend may use: they read
// A from the key store.
size_t kem_workspace_bytes(const KemParams &P, int64_t B, bool with_A = true);
void kem_workspace_carve(const KemParams &P, int64_t B, uint8_t *base, KemWorkspace &W, bool with_A = true);

// backend: 0 = MFMA, otherwise VALU.  seeds: [B][96] = seed_A | seed_s | z.
hipError_t kem_keypair(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *seeds, uint8_t *pk,
                       uint8_t *sk, int64_t B, int *status, hipStream_t st);
// rnd: [B][32] randomness per instance.
hipError_t kem_encaps(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *pk, const uint8_t *rnd,
                      uint8_t *ct, uint8_t *ss, int64_t B, int *status, hipStream_t st);
// fail (optional): [B] bytes, 1 where the ciphertext was rejected.
hipError_t kem_decaps(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *ct, const uint8_t *sk,
                      uint8_t *ss, uint8_t *fail, int64_t B, int *status, hipStream_t st);

// Encapsulation / decapsulation for instance t under key keys[t] of K
// (ids checked by the caller); A and H(pk) come from the store.
hipError_t kem_encaps_keyed(const KemParams &P, const KemWorkspace &W, int backend, const KeyStore &K,
                            const int32_t *keys, const uint8_t *rnd, uint8_t *ct, uint8_t *ss, int64_t B, int *status,
                            hipStream_t st);
hipError_t kem_decaps_keyed(const KemParams &P, const KemWorkspace &W, int backend, const KeyStore &K,
                            const int32_t *keys, const uint8_t *ct, uint8_t *ss, uint8_t *fail, int64_t B,
                            int *status, hipStream_t st);
// A and H(pk) of n public keys.
hipError_t kem_store_derive(const KemParams &P, const uint8_t *pk, int64_t n, uint16_t *A, uint8_t *hpk,
                            hipStream_t st);

} // namespace tmvp
