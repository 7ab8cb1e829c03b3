/*
 * The JNI entry points libfgkl.so exports for the reference's Java bindings
 * (src/main/java/com/fulcrumgenomics/fgkl/pairhmm/FgklPairHmm.java:98-106,
 *  src/main/java/com/fulcrumgenomics/fgkl/pdhmm/FgklPdHmm.java:113-121,
 *  src/main/java/com/fulcrumgenomics/fgkl/smithwaterman/FgklSmithWaterman.java:45-66).
 * Declared with opaque pointers (JNIEnv*, jclass, arrays) because the engine is built without
 * jni.h; the calling convention is the platform C one, as JNIEXPORT/JNICALL are on Linux.
 */
#ifndef FGKL_JNI_H
#define FGKL_JNI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces crates/jni/src/lib.rs:149-156: returns a jstring naming the engine. */
void* Java_com_fulcrumgenomics_fgkl_pairhmm_FgklPairHmm_backendNative(void* env, void* cls);

/* Replaces crates/jni/src/lib.rs:158-195 (FgklPairHmm.computeNative). */
void Java_com_fulcrumgenomics_fgkl_pairhmm_FgklPairHmm_computeNative(
    void* env, void* cls, void* bases, void* quals, void* ins_gop, void* del_gop, void* gcp,
    void* read_offsets, void* hap_bases, void* hap_offsets, uint8_t double_precision,
    void* likelihoods);

/* Replaces crates/jni/src/lib.rs:197-204. */
void* Java_com_fulcrumgenomics_fgkl_pdhmm_FgklPdHmm_backendNative(void* env, void* cls);

/* Replaces crates/jni/src/lib.rs:206-254 (FgklPdHmm.computeNative). */
void Java_com_fulcrumgenomics_fgkl_pdhmm_FgklPdHmm_computeNative(
    void* env, void* cls, void* bases, void* quals, void* ins_gop, void* del_gop, void* gcp,
    void* read_offsets, void* hap_bases, void* hap_flags, void* hap_offsets,
    uint8_t double_precision, void* likelihoods);

/* Replaces crates/jni/src/lib.rs:256-268: returns a jstring naming the aligner's engine. */
void* Java_com_fulcrumgenomics_fgkl_smithwaterman_FgklSmithWaterman_backendNative(void* env,
                                                                                 void* cls);

/* Replaces crates/jni/src/lib.rs:271-313 (FgklSmithWaterman.alignNative): returns the CIGAR as
 * a jstring and writes the alignment offset into offset[0]. */
void* Java_com_fulcrumgenomics_fgkl_smithwaterman_FgklSmithWaterman_alignNative(
    void* env, void* cls, void* reference, void* alternate, int32_t match_value,
    int32_t mismatch_penalty, int32_t gap_open, int32_t gap_extend, int32_t strategy,
    void* offset);

#ifdef __cplusplus
}
#endif
#endif /* FGKL_JNI_H */
