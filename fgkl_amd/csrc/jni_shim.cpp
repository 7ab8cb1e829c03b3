// JNI entry points of FgklPairHmm, FgklPdHmm and FgklSmithWaterman (crates/jni/src/lib.rs:149-313),
// exported from libfgkl.so so the reference's Java bindings load this engine unchanged
// (NativeLoader.java).
//
// No jni.h is available here, so the few JNI types and the function-table slots used are
// declared below; the slot numbers are the JNI specification's fixed interface-table indices.
// Inputs are copied out of the JVM before any computation, no critical section is held, and the
// output is written with one region call (lib.rs:5-7, :191).  Validation failures throw
// IllegalArgumentException, other failures RuntimeException; an exception already pending is
// never overwritten (lib.rs:41-74).
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fgkl.h"

namespace {

using jint = int32_t;
using jsize = jint;
using jboolean = uint8_t;
using jbyte = int8_t;
using jdouble = double;
using jobject = void*;
using jclass = jobject;
using jstring = jobject;
using jarray = jobject;
using jbyteArray = jarray;
using jintArray = jarray;
using jdoubleArray = jarray;
using jthrowable = jobject;

// The JNI interface function table: a fixed array of function pointers.
struct JNINativeInterface {
  void* slots[233];
};
using JNIEnv = const JNINativeInterface*;

constexpr int kFindClass = 6;
constexpr int kThrowNew = 14;
constexpr int kNewStringUTF = 167;
constexpr int kGetArrayLength = 171;
constexpr int kGetByteArrayRegion = 200;
constexpr int kGetIntArrayRegion = 203;
constexpr int kSetIntArrayRegion = 211;
constexpr int kSetDoubleArrayRegion = 214;
constexpr int kExceptionCheck = 228;

template <typename F>
F fn(JNIEnv* env, int slot) {
  return reinterpret_cast<F>((*env)->slots[slot]);
}

bool pending(JNIEnv* env) {
  return fn<jboolean (*)(JNIEnv*)>(env, kExceptionCheck)(env) != 0;
}

void throw_new(JNIEnv* env, const char* cls, const std::string& msg) {
  if (pending(env)) return;
  jclass c = fn<jclass (*)(JNIEnv*, const char*)>(env, kFindClass)(env, cls);
  if (!c) return;  // FindClass left its own exception pending
  fn<jint (*)(JNIEnv*, jclass, const char*)>(env, kThrowNew)(env, c, msg.c_str());
}

const char* const kIllegalArgument = "java/lang/IllegalArgumentException";
const char* const kRuntimeException = "java/lang/RuntimeException";
const char* const kIndexOutOfBounds = "java/lang/ArrayIndexOutOfBoundsException";

// ReadBatch::fetch's copies (lib.rs:105-133): a null array is a JNI error.
bool copy_bytes(JNIEnv* env, jbyteArray a, std::vector<uint8_t>& out) {
  if (!a) {
    throw_new(env, kRuntimeException, "JNI error: null array");
    return false;
  }
  const jsize n = fn<jsize (*)(JNIEnv*, jarray)>(env, kGetArrayLength)(env, a);
  out.resize(size_t(n));
  if (n)
    fn<void (*)(JNIEnv*, jbyteArray, jsize, jsize, jbyte*)>(env, kGetByteArrayRegion)(
        env, a, 0, n, reinterpret_cast<jbyte*>(out.data()));
  return !pending(env);
}

bool copy_ints(JNIEnv* env, jintArray a, std::vector<int32_t>& out) {
  if (!a) {
    throw_new(env, kRuntimeException, "JNI error: null array");
    return false;
  }
  const jsize n = fn<jsize (*)(JNIEnv*, jarray)>(env, kGetArrayLength)(env, a);
  out.resize(size_t(n));
  if (n)
    fn<void (*)(JNIEnv*, jintArray, jsize, jsize, jint*)>(env, kGetIntArrayRegion)(env, a, 0, n,
                                                                                   out.data());
  return !pending(env);
}

void compute(JNIEnv* env, bool pd, jbyteArray bases, jbyteArray quals, jbyteArray ins,
             jbyteArray del, jbyteArray gcp, jintArray read_offsets, jbyteArray hap_bases,
             jbyteArray hap_flags, jintArray hap_offsets, jboolean double_precision,
             jdoubleArray likelihoods) {
  std::vector<uint8_t> b, q, i, d, g, hb, hf;
  std::vector<int32_t> ro, ho;
  if (!copy_bytes(env, bases, b) || !copy_bytes(env, quals, q) || !copy_bytes(env, ins, i) ||
      !copy_bytes(env, del, d) || !copy_bytes(env, gcp, g) || !copy_ints(env, read_offsets, ro) ||
      !copy_bytes(env, hap_bases, hb) || (pd && !copy_bytes(env, hap_flags, hf)) ||
      !copy_ints(env, hap_offsets, ho))
    return;
  if (!likelihoods) {
    throw_new(env, kRuntimeException, "JNI error: null array");
    return;
  }
  const jsize n_out = fn<jsize (*)(JNIEnv*, jarray)>(env, kGetArrayLength)(env, likelihoods);
  std::vector<double> out(static_cast<size_t>(n_out));
  const int rc =
      pd ? fgkl_pdhmm_compute_native(b.data(), int64_t(b.size()), q.data(), int64_t(q.size()),
                                     i.data(), int64_t(i.size()), d.data(), int64_t(d.size()),
                                     g.data(), int64_t(g.size()), ro.data(), int64_t(ro.size()),
                                     hb.data(), int64_t(hb.size()), hf.data(), int64_t(hf.size()),
                                     ho.data(), int64_t(ho.size()), double_precision != 0,
                                     out.data(), int64_t(out.size()))
         : fgkl_pairhmm_compute_native(b.data(), int64_t(b.size()), q.data(), int64_t(q.size()),
                                       i.data(), int64_t(i.size()), d.data(), int64_t(d.size()),
                                       g.data(), int64_t(g.size()), ro.data(), int64_t(ro.size()),
                                       hb.data(), int64_t(hb.size()), ho.data(),
                                       int64_t(ho.size()), double_precision != 0, out.data(),
                                       int64_t(out.size()));
  if (rc == FGKL_INVALID_ARGUMENT) {
    throw_new(env, kIllegalArgument, fgkl_last_error());
    return;
  }
  if (rc != FGKL_OK) {
    throw_new(env, kRuntimeException, std::string("fgkl native code failed: ") + fgkl_last_error());
    return;
  }
  if (n_out)
    fn<void (*)(JNIEnv*, jdoubleArray, jsize, jsize, const jdouble*)>(env, kSetDoubleArrayRegion)(
        env, likelihoods, 0, n_out, out.data());
}

jstring backend(JNIEnv* env) {
  return fn<jstring (*)(JNIEnv*, const char*)>(env, kNewStringUTF)(env, fgkl_backend_name());
}

// crates/jni/src/lib.rs:282-312: the strategy code is checked first, then the arrays are copied,
// the pair aligned, the offset written into offset[0] and the CIGAR returned as a string.
jstring sw_align(JNIEnv* env, jbyteArray reference, jbyteArray alternate, jint match_value,
                 jint mismatch, jint gap_open, jint gap_extend, jint strategy, jintArray offset) {
  if (strategy < 0 || strategy > 3) {
    throw_new(env, kIllegalArgument, "unknown overhang strategy code " + std::to_string(strategy));
    return nullptr;
  }
  std::vector<uint8_t> r, a;
  if (!copy_bytes(env, reference, r) || !copy_bytes(env, alternate, a)) return nullptr;
  std::vector<char> text(2 * (r.size() + a.size()) + 64);
  int32_t off = 0;
  const int rc = fgkl_sw_align_native(r.data(), int64_t(r.size()), a.data(), int64_t(a.size()),
                                      match_value, mismatch, gap_open, gap_extend, strategy, &off,
                                      text.data(), int64_t(text.size()));
  if (rc == FGKL_INVALID_ARGUMENT) {
    throw_new(env, kIllegalArgument, fgkl_last_error());
    return nullptr;
  }
  if (rc != FGKL_OK) {
    throw_new(env, kRuntimeException, std::string("fgkl native code failed: ") + fgkl_last_error());
    return nullptr;
  }
  if (!offset) {
    throw_new(env, kRuntimeException, "JNI error: null array");
    return nullptr;
  }
  if (fn<jsize (*)(JNIEnv*, jarray)>(env, kGetArrayLength)(env, offset) < 1) {
    throw_new(env, kIndexOutOfBounds, "offset array is empty");
    return nullptr;
  }
  fn<void (*)(JNIEnv*, jintArray, jsize, jsize, const jint*)>(env, kSetIntArrayRegion)(env, offset, 0, 1,
                                                                                       &off);
  if (pending(env)) return nullptr;
  return fn<jstring (*)(JNIEnv*, const char*)>(env, kNewStringUTF)(env, text.data());
}

}  // namespace

extern "C" {

// crates/jni/src/lib.rs:149-156
__attribute__((visibility("default"))) jstring
Java_com_fulcrumgenomics_fgkl_pairhmm_FgklPairHmm_backendNative(JNIEnv* env, jclass) {
  return backend(env);
}

// crates/jni/src/lib.rs:158-195
__attribute__((visibility("default"))) void
Java_com_fulcrumgenomics_fgkl_pairhmm_FgklPairHmm_computeNative(
    JNIEnv* env, jclass, jbyteArray bases, jbyteArray quals, jbyteArray ins_gop,
    jbyteArray del_gop, jbyteArray gcp, jintArray read_offsets, jbyteArray hap_bases,
    jintArray hap_offsets, jboolean double_precision, jdoubleArray likelihoods) {
  try {
    compute(env, false, bases, quals, ins_gop, del_gop, gcp, read_offsets, hap_bases, nullptr,
            hap_offsets, double_precision, likelihoods);
  } catch (const std::exception& e) {
    throw_new(env, kRuntimeException, std::string("fgkl native code panicked: ") + e.what());
  }
}

// crates/jni/src/lib.rs:197-204
__attribute__((visibility("default"))) jstring
Java_com_fulcrumgenomics_fgkl_pdhmm_FgklPdHmm_backendNative(JNIEnv* env, jclass) {
  return backend(env);
}

// crates/jni/src/lib.rs:206-254
__attribute__((visibility("default"))) void
Java_com_fulcrumgenomics_fgkl_pdhmm_FgklPdHmm_computeNative(
    JNIEnv* env, jclass, jbyteArray bases, jbyteArray quals, jbyteArray ins_gop,
    jbyteArray del_gop, jbyteArray gcp, jintArray read_offsets, jbyteArray hap_bases,
    jbyteArray hap_flags, jintArray hap_offsets, jboolean double_precision,
    jdoubleArray likelihoods) {
  try {
    compute(env, true, bases, quals, ins_gop, del_gop, gcp, read_offsets, hap_bases, hap_flags,
            hap_offsets, double_precision, likelihoods);
  } catch (const std::exception& e) {
    throw_new(env, kRuntimeException, std::string("fgkl native code panicked: ") + e.what());
  }
}

// crates/jni/src/lib.rs:256-268
__attribute__((visibility("default"))) jstring
Java_com_fulcrumgenomics_fgkl_smithwaterman_FgklSmithWaterman_backendNative(JNIEnv* env, jclass) {
  return fn<jstring (*)(JNIEnv*, const char*)>(env, kNewStringUTF)(env, fgkl_sw_backend_name());
}

// crates/jni/src/lib.rs:271-313
__attribute__((visibility("default"))) jstring
Java_com_fulcrumgenomics_fgkl_smithwaterman_FgklSmithWaterman_alignNative(
    JNIEnv* env, jclass, jbyteArray reference, jbyteArray alternate, jint match_value,
    jint mismatch_penalty, jint gap_open, jint gap_extend, jint strategy, jintArray offset) {
  try {
    return sw_align(env, reference, alternate, match_value, mismatch_penalty, gap_open, gap_extend,
                    strategy, offset);
  } catch (const std::exception& e) {
    throw_new(env, kRuntimeException, std::string("fgkl native code panicked: ") + e.what());
    return nullptr;
  }
}

}  // extern "C"
