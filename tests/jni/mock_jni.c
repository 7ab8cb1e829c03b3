/*
 * A mock JNIEnv for testing libfgkl.so's JNI entry points without a JVM (test infrastructure).
 * Java arrays are {kind, length, data} records; the interface table fills the slots the shim
 * uses (FindClass 6, ThrowNew 14, NewStringUTF 167, GetArrayLength 171, GetByteArrayRegion 200,
 * GetIntArrayRegion 203, SetIntArrayRegion 211, SetDoubleArrayRegion 214, ExceptionCheck 228) and
 * records exceptions.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int kind; int32_t len; void* data; } jarr;
typedef struct { void* slots[233]; } table_t;

static int g_pending;
static char g_class[256], g_msg[2048], g_str[4096];
static int g_set_calls;

static uint8_t ExceptionCheck(void* env) { (void)env; return (uint8_t)g_pending; }
static void* FindClass(void* env, const char* name) {
  (void)env;
  strncpy(g_class, name, sizeof(g_class) - 1);
  return (void*)g_class;
}
static int32_t ThrowNew(void* env, void* cls, const char* msg) {
  (void)env; (void)cls;
  g_pending = 1;
  strncpy(g_msg, msg, sizeof(g_msg) - 1);
  return 0;
}
static void* NewStringUTF(void* env, const char* s) { (void)env; strncpy(g_str, s, sizeof(g_str) - 1); return g_str; }
static int32_t GetArrayLength(void* env, void* a) { (void)env; return ((jarr*)a)->len; }
static void GetByteArrayRegion(void* env, void* a, int32_t s, int32_t n, int8_t* buf) {
  (void)env; memcpy(buf, (int8_t*)((jarr*)a)->data + s, (size_t)n);
}
static void GetIntArrayRegion(void* env, void* a, int32_t s, int32_t n, int32_t* buf) {
  (void)env; memcpy(buf, (int32_t*)((jarr*)a)->data + s, sizeof(int32_t) * (size_t)n);
}
static void SetIntArrayRegion(void* env, void* a, int32_t s, int32_t n, const int32_t* buf) {
  (void)env; g_set_calls++; memcpy((int32_t*)((jarr*)a)->data + s, buf, sizeof(int32_t) * (size_t)n);
}
static void SetDoubleArrayRegion(void* env, void* a, int32_t s, int32_t n, const double* buf) {
  (void)env; g_set_calls++; memcpy((double*)((jarr*)a)->data + s, buf, sizeof(double) * (size_t)n);
}

static table_t g_table;
static const table_t* g_env = &g_table;

static void init(void) {
  memset(&g_table, 0, sizeof(g_table));
  g_table.slots[6] = (void*)FindClass;
  g_table.slots[14] = (void*)ThrowNew;
  g_table.slots[167] = (void*)NewStringUTF;
  g_table.slots[171] = (void*)GetArrayLength;
  g_table.slots[200] = (void*)GetByteArrayRegion;
  g_table.slots[203] = (void*)GetIntArrayRegion;
  g_table.slots[211] = (void*)SetIntArrayRegion;
  g_table.slots[214] = (void*)SetDoubleArrayRegion;
  g_table.slots[228] = (void*)ExceptionCheck;
  g_pending = 0; g_class[0] = 0; g_msg[0] = 0; g_str[0] = 0; g_set_calls = 0;
}

typedef void (*compute_fn)(const void*, void*, void*, void*, void*, void*, void*, void*, void*,
                           void*, uint8_t, void*);
typedef void (*pd_compute_fn)(const void*, void*, void*, void*, void*, void*, void*, void*, void*,
                              void*, void*, uint8_t, void*);
typedef void* (*backend_fn)(const void*, void*);

/* Calls a computeNative entry with Java-like arrays; `nulls` has bit i set to pass null for
 * argument i (0 bases .. 4 gcp, 5 read offsets, 6 hap bases, 7 hap flags, 8 hap offsets,
 * 9 likelihoods).  Returns 1 if an exception is pending afterwards. */
int mock_compute(void* fnp, int pd, int8_t* bases, int32_t nb, int8_t* quals, int32_t nq,
                 int8_t* ins, int32_t ni, int8_t* del, int32_t nd, int8_t* gcp, int32_t ng,
                 int32_t* ro, int32_t nro, int8_t* hb, int32_t nhb, int8_t* hf, int32_t nhf,
                 int32_t* ho, int32_t nho, int dp, double* out, int32_t nout, int nulls) {
  init();
  jarr a[10] = {{1, nb, bases}, {1, nq, quals}, {1, ni, ins}, {1, nd, del}, {1, ng, gcp},
                {2, nro, ro}, {1, nhb, hb}, {1, nhf, hf}, {2, nho, ho}, {3, nout, out}};
  void* p[10];
  for (int i = 0; i < 10; ++i) p[i] = (nulls >> i) & 1 ? NULL : (void*)&a[i];
  if (pd)
    ((pd_compute_fn)fnp)(&g_env, NULL, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8],
                         (uint8_t)dp, p[9]);
  else
    ((compute_fn)fnp)(&g_env, NULL, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[8], (uint8_t)dp,
                      p[9]);
  return g_pending;
}

typedef void* (*sw_align_fn)(const void*, void*, void*, void*, int32_t, int32_t, int32_t, int32_t,
                             int32_t, void*);

/* Calls FgklSmithWaterman.alignNative with Java-like arrays (`nulls` bit 0 reference, 1
 * alternate, 2 offset; n_offset is the offset array's length).  Returns the CIGAR string or NULL
 * (then an exception is pending). */
const char* mock_sw_align(void* fnp, int8_t* ref, int32_t nref, int8_t* alt, int32_t nalt,
                          const int32_t* params, int32_t strategy, int32_t* offset,
                          int32_t n_offset, int nulls) {
  init();
  jarr a[3] = {{1, nref, ref}, {1, nalt, alt}, {2, n_offset, offset}};
  void* p[3];
  for (int i = 0; i < 3; ++i) p[i] = (nulls >> i) & 1 ? NULL : (void*)&a[i];
  void* s = ((sw_align_fn)fnp)(&g_env, NULL, p[0], p[1], params[0], params[1], params[2],
                               params[3], strategy, p[2]);
  return s ? (const char*)s : NULL;
}

const char* mock_backend(void* fnp) {
  init();
  ((backend_fn)fnp)(&g_env, NULL);
  return g_str;
}

const char* mock_exception_class(void) { return g_class; }
const char* mock_exception_message(void) { return g_msg; }
int mock_set_region_calls(void) { return g_set_calls; }
