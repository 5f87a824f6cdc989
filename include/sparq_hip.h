/*
 * sparq_hip.h -- thin host->HIP C ABI of the MI355X sparse-state engine.
 *
 * The engine keeps a sparse quantum state as a structure of arrays on one GPU:
 * W packed uint64 key words per entry (word 0 most significant, word-major in
 * HBM) and one FP64 complex amplitude per entry.  Registers are bit fields of
 * the key (sparq_field_t); the C++ host layer (sparqsim_amd/csrc/host) maps the
 * reference's process-global register table onto fields so that unsigned
 * comparison of the packed words equals the reference canonical order
 * System::operator< (SparQ/src/basic_components.cpp:461-474).
 *
 * Every call returns a status code; no C++ exception crosses this boundary.
 * All work for one state is ordered on that state's HIP stream.  Calls that
 * report a size (sparq_state_size, ...) synchronize that stream.
 *
 * The reference has no counterpart of this layer (its CUDA path calls thrust
 * directly in SparQ/src/cuda); each entry point below names the reference
 * operator(s) whose semantics it implements.
 */
#ifndef SPARQ_HIP_H
#define SPARQ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
enum sparq_status {
    SPARQ_OK = 0,
    SPARQ_OVERFLOW = 1,   /* a predicted capacity was too small; the exact path must be rerun */
    SPARQ_OOM = 2,        /* device allocation failed; the state is unchanged */
    SPARQ_INVALID = 3,    /* bad argument (maps to std::invalid_argument / ValueError) */
    SPARQ_RUNTIME = 4,    /* HIP error or violated invariant (std::runtime_error / RuntimeError) */
    SPARQ_NO_DEVICE = 5   /* no usable GPU */
};

#define SPARQ_MAX_WORDS 8          /* 512 key bits */
#define SPARQ_MAX_PROGRAM_OPS 16   /* ops fused into one pointwise launch */
#define SPARQ_MAX_PROGRAM_CONDS 32 /* condition predicates per pointwise launch */

typedef struct sparq_state_s* sparq_state_t;
typedef struct sparq_table_s* sparq_table_t;

/* A register inside the packed key: bits [shift, shift+width) of key word `word`. */
typedef struct {
    uint8_t word;
    uint8_t shift;
    uint8_t width;
    uint8_t flags; /* bit 0: register is SignedInteger (used by sign-extending ops) */
} sparq_field_t;

/* Condition predicate, AND-combined (global_macros.h:28-92). */
enum sparq_cond_kind {
    SPARQ_COND_NONZERO = 0,  /* field != 0                         (conditioned_by_nonzeros) */
    SPARQ_COND_ALL_ONES = 1, /* field == 2^width - 1               (conditioned_by_all_ones) */
    SPARQ_COND_BIT = 2,      /* bit `value` of field is 1          (conditioned_by_bit)      */
    SPARQ_COND_VALUE = 3,    /* field == value                     (conditioned_by_value)    */
    SPARQ_COND_GE = 4,       /* field >= value      (PartialTraceSelectRange lower bound, partial_trace.cpp:187-225) */
    SPARQ_COND_LE = 5,       /* field <= value      (PartialTraceSelectRange upper bound) */
    SPARQ_COND_BIT_CLEAR = 6 /* bit `value` of field is 0: a control on |0> (the reference
                                reaches it by flipping the bit around a conditioned_by_bit
                                operator, PySparQ/pysparq/rir.py _with_conditions) */
};
typedef struct {
    uint32_t kind;
    sparq_field_t field;
    uint64_t value;
} sparq_cond_t;

/* Per-entry operations (key rewrites and diagonal phases).  Operand fields are
 * f[0..3]; integer constants c[0..2]; complex constants d[0..3] as (re,im)
 * pairs.  Each op applies only to entries satisfying its conditions
 * conds[cond_first .. cond_first+ncond). */
enum sparq_pw_opcode {
    SPARQ_PW_FLIP_BIT = 1,        /* f0 ^= 1<<c0                                  X_Bool (basic_gates.cpp:383-400) */
    SPARQ_PW_FLIP_ALL,            /* f0 = ~f0                                     FlipBools (quantum_arithmetic.cpp:11-30) */
    SPARQ_PW_SWAP_BITS,           /* swap bit c0 of f0 with bit c1 of f1          Swap_Bool_Bool (:32-61) */
    SPARQ_PW_SWAP_FIELDS,         /* swap f0, f1                                  Swap_General_General (:982-999) */
    SPARQ_PW_ROTL,                /* rotate f0 left by c0                         ShiftLeft_InPlace (:63-93) */
    SPARQ_PW_ROTR,                /* rotate f0 right by c0                        ShiftRight_InPlace (:95-121) */
    SPARQ_PW_XOR_MUL_CONST,       /* f1 ^= f0*c0                                  Mult_UInt_ConstUInt (:137-155) */
    SPARQ_PW_ADD_MUL_CONST,       /* f1 += f0*c0  (c0 two's complement for dag)   Add_Mult_UInt_ConstUInt_InPlace (:157-195) */
    SPARQ_PW_MODMUL_CONST,        /* f0 = (f0*c0) % c1                            Mod_Mult_UInt_ConstUInt_InPlace (:222-285) */
    SPARQ_PW_XOR_ADD,             /* f2 ^= f0+f1                                  Add_UInt_UInt (:287-306) */
    SPARQ_PW_ADD_FIELD,           /* f1 += f0                                     Add_UInt_UInt_InPlace (:308-327) */
    SPARQ_PW_SUB_FIELD,           /* f1 -= f0                                     Add_UInt_UInt_InPlace::dag (:329-349) */
    SPARQ_PW_XOR_ADD_CONST,       /* f1 ^= f0+c0                                  Add_UInt_ConstUInt (:351-369) */
    SPARQ_PW_ADD_CONST,           /* f0 += c0                                     Add_ConstUInt_InPlace (:371-411) */
    SPARQ_PW_XOR_SUB,             /* f2 ^= f0-f1                                  Sub_UInt_UInt (:490-509) */
    SPARQ_PW_XOR_NEG,             /* f1 ^= -f0                                    Neg_UInt (:511-530) */
    SPARQ_PW_XOR_ABS,             /* f1 ^= |sext(f0)|                             Abs_SInt (:532-554) */
    SPARQ_PW_XOR_MUL,             /* f2 ^= f0*f1                                  Mul_UInt_UInt (:556-575) */
    SPARQ_PW_XOR_DIV,             /* f2 ^= f1 ? f0/f1 : 0                         Div_UInt_UInt (:577-600) */
    SPARQ_PW_XOR_ISQRT,           /* f1 ^= isqrt(f0)                              Sqrt_UInt (:602-621) */
    SPARQ_PW_XOR_SELECT,          /* f3 ^= (f0&1) ? f1 : f2                       Select_Bool_UInt_UInt (:623-644) */
    SPARQ_PW_XOR_AND,             /* f2 ^= f0&f1                                  And_UInt_UInt (:646-665) */
    SPARQ_PW_XOR_OR,              /* f2 ^= f0|f1                                  Or_UInt_UInt (:667-686) */
    SPARQ_PW_XOR_XOR,             /* f2 ^= f0^f1                                  Xor_UInt_UInt (:688-707) */
    SPARQ_PW_FLAG_LESS_S,         /* f2 ^= sext(f0) < sext(f1)                    Less_SInt_SInt (:709-729) */
    SPARQ_PW_FLAG_CARRY,          /* f3 ^= carry(f0+f1) at width(f2)              Carry_UInt_UInt (:731-754) */
    SPARQ_PW_FLAG_OVERFLOW_S,     /* f3 ^= signed overflow at width(f2)           Overflow_SInt_SInt (:756-785) */
    SPARQ_PW_FLAG_MULOVERFLOW,    /* f3 ^= f0*f1 overflows width(f2)              MulOverflow_UInt_UInt (:787-817) */
    SPARQ_PW_FLAG_ISZERO,         /* f1 ^= f0 == 0                                IsZero_UInt (:819-837) */
    SPARQ_PW_FLAG_NEGATIVE,       /* f1 ^= sext(f0) < 0                           Negative_SInt (:839-856) */
    SPARQ_PW_ADD_ANY,             /* f0 += ext(f1)                                Add_AnyInt_AnyInt_InPlace (:869-887) */
    SPARQ_PW_SUB_ANY,             /* f0 -= ext(f1)                                Add_AnyInt_AnyInt_InPlace::dag (:889-907) */
    SPARQ_PW_XOR_ASSIGN,          /* f1 ^= f0                                     Assign (:909-930) */
    SPARQ_PW_COMPARE,             /* f0==f1 ? f3^=1 : f2^=(f0<f1)                 Compare_UInt_UInt (:932-959) */
    SPARQ_PW_FLAG_LESS_U,         /* f2 ^= f0 < f1                                Less_UInt_UInt (:961-980) */
    SPARQ_PW_XOR_MID,             /* f2 ^= (f0+f1)/2                              GetMid_UInt_UInt (:1001-1020) */
    SPARQ_PW_XOR_DIVSQRT_ACOS,    /* f2 ^= rational(acos(sqrt(f0/f1))/2pi)        Div_Sqrt_Arccos_UInt_UInt (:413-433) */
    SPARQ_PW_XOR_SQRTDIV_ACOS,    /* f2 ^= rational(acos(sext(f0)/sqrt(f1))/2pi)  Sqrt_Div_Arccos_Int_UInt (:435-456) */
    SPARQ_PW_XOR_ROTATE_ANGLE,    /* f2 ^= rational(atan2(sext f1, sext f0)/2pi)  GetRotateAngle_Int_Int (:458-488) */
    SPARQ_PW_QRAM_LOAD,           /* f1 ^= table[f0 mod 2^c0]                     QRAMLoad::noise_free_impl (qram.cpp:15-35) */
    SPARQ_PW_PHASE_BITS,          /* for each set bit b of (f0 & c0), ascending:
                                     amp *= d0                                    Phase_Bool/Z/S/T (basic_gates.cpp:11-55) */
    SPARQ_PW_RZ,                  /* bit c0 of f0 ? amp*=d0 : amp*=d1             RZ_Bool (basic_gates.cpp:486-520) */
    SPARQ_PW_Y,                   /* flip bit c0; bit set ? amp*=i : amp*=-i      Y_Bool (basic_gates.cpp:402-420) */
    SPARQ_PW_ROT_DIAG,            /* bit c0 of f0 ? amp*=d1 : amp*=d0            Rot_Bool diagonal (basic_gates.cpp:104-122) */
    SPARQ_PW_ROT_OFFDIAG,         /* bit c0 set ? amp*=d1 : amp*=d0; flip bit     Rot_Bool off-diagonal (basic_gates.cpp:134-157) */
    SPARQ_PW_NEG_AMP,             /* amp *= -1                                    ZeroConditionalPhaseFlip / Reflection_Bool
                                                                                  (parallel_phase_operations.cpp:11-70) */
    SPARQ_PW_MUL_AMP,             /* amp *= d0                                    GlobalPhase (parallel_phase_operations.cpp:72-108) */
    SPARQ_PW_SET_FIELD,           /* f0 = c0                                      Init_Unsafe (dark_magic.cpp:35-41) */
    SPARQ_PW_INC_WRAP,            /* f0 += 1 mod 2^w; on wrap f1 = (f1+1)%2       PlusOneAndOverflow
                                     (SparQ_Algorithm/src/BlockEncoding/block_encoding_tridiagonal.cpp:20-47) */
    SPARQ_PW_DEC_WRAP,            /* f0 -= 1 mod 2^w; on wrap f1 = (f1+1)%2       PlusOneAndOverflow::dag (:55-84) */
    /* Sparse-matrix access of the CKS quantum walk (SparQ_Algorithm/src/hamiltonian_simulation.cpp) */
    SPARQ_PW_XOR_AFFINE,          /* f3 ^= f0 + c0*f1 (+ f2 when f2.width)        GetRowAddr (:222-237), GetDataAddr (:239-255) */
    SPARQ_PW_XOR_BSEARCH,         /* f2 ^= search of f1 in table[f0, f0+c0):      QuantumBinarySearch_Fast (:182-220)
                                     c1 halvings, mid = (l+r)/2, hit -> mid,
                                     miss -> 0; table index masked by c2 */
    SPARQ_PW_XOR_QW_ANGLE,        /* f1 ^= rational(acos(sqrt(r))/2pi), r =       GetQWRotateAngle_Int_Int_Int (:257-290)
                                     |f0| / Amax on c0 data bits; c1 = 1:
                                     unsigned, Amax = 2^c0-1; c1 = 0: two's
                                     complement, Amax = 2^(c0-1)-1 */
    /* Direct oracles of the tridiagonal walk (SparQ_Algorithm/src/hamiltonian_tridiagonal.cpp); the
       matrix is c0 = n_row | k_width << 48, c1 = alpha_enc, c2 = beta_enc; fields f0 offset, f1 row,
       f2 slot / column, f3 output / search result.  M(i) is TridiagonalSpec::combined_mem
       (SparQ_Algorithm/include/hamiltonian_simulation_tridiagonal.h:104-190). */
    SPARQ_PW_TRIDIAG_LOAD,        /* f3 ^= M(f0 + 4*f1 + f2)                      TridiagOracle1 (:16-34) */
    SPARQ_PW_TRIDIAG_COL2SLOT,    /* column -> slot                               TridiagOracle2 (:36-61) */
    SPARQ_PW_TRIDIAG_SLOT2COL,    /* slot -> column                               TridiagOracle2::dag (:63-88) */
    /* No counterpart in the reference: */
    SPARQ_PW_ASSERT_ZERO,         /* f0 must be 0 (else the capture's flags get kCapNonzeroRemoved); f0 = 0.
                                     RemoveRegister's uniformity check (system_operations.cpp:303-322) inside a
                                     captured sequence; the engine supplies the flags word, only valid there */
    SPARQ_PW_PAULI_TABLE,         /* seeded Pauli noise from a device table of uint64 words: k = table[0]
                                     (iteration), s = f0 (segment, e.g. trajectory id); when k < c1 and
                                     s < c0: f1 ^= table[1 + 2(k*c0 + s)], then amp *= d0 for each set bit of
                                     f1 & table[2 + 2(k*c0 + s)], ascending -- the X_Bool / Z_Bool insertions
                                     of one iteration of one segment (basic_gates.cpp:11-55, 383-400) */
    SPARQ_PW_OPCODE_END
};

typedef struct {
    uint32_t opcode;
    uint16_t cond_first;
    uint16_t ncond;
    sparq_field_t f[4];
    uint64_t c[3];
    double d[8];            /* d[2k], d[2k+1] = re, im of complex constant k */
    const void* table;      /* device table (sparq_table_device_ptr) for SPARQ_PW_QRAM_LOAD */
    uint32_t table_word_bytes;
    uint32_t pad;
} sparq_pw_op_t;

/* ------------------------------------------------------------ device & state */
int sparq_device_init(int device);
int sparq_device_name(char* buf, size_t len);

/* Create an empty state with `words` key words.  (SparseState(), basic_components.h:611-624,
 * is sparq_state_create + sparq_state_set_basis(0).) */
int sparq_state_create(uint32_t words, uint64_t capacity, sparq_state_t* out);
int sparq_state_destroy(sparq_state_t st);
int sparq_state_clone(sparq_state_t src, sparq_state_t* out);
int sparq_state_words(sparq_state_t st, uint32_t* words);
int sparq_state_size(sparq_state_t st, uint64_t* n);
/* Replace the content by the single entry |key> with amplitude 1. */
int sparq_state_set_basis(sparq_state_t st, const uint64_t* key);
/* keys: n*W words, entry-major (entry i word w at keys[i*W+w]); amps: 2n doubles (re,im). */
int sparq_state_upload(sparq_state_t st, const uint64_t* keys, const double* amps, uint64_t n);
int sparq_state_download(sparq_state_t st, uint64_t* keys, double* amps, uint64_t capacity, uint64_t* n);
/* Move fields to a new layout (register add/remove/split/combine, system_operations.cpp:124-369).
 * Bits of the new key not covered by any destination field are zero. */
int sparq_state_relayout(sparq_state_t st, uint32_t new_words,
                         const sparq_field_t* from, const sparq_field_t* to, uint32_t nfields);
/* Change the key width without moving fields: words below min(old, new) keep
 * their content, added words read as zero, removed words are dropped (the caller
 * guarantees they hold no active register).  Growing back to a width the state
 * held before reuses its allocation. */
int sparq_state_resize_words(sparq_state_t st, uint32_t new_words);
int sparq_state_sync(sparq_state_t st);
/* Raw device pointers of the current arrays (valid until the next call that changes the size).
   Once called, set_basis no longer marks the state's amplitude as known (the caller may write it). */
int sparq_state_device_arrays(sparq_state_t st, uint64_t** keys, double** amps, uint64_t* capacity);
/* The HIP stream (hipStream_t) that all work of this state is ordered on. */
int sparq_state_stream(sparq_state_t st, void** stream);

/* Device lookup table for QRAMLoad (the QRAM memory, qram.cpp:29-31). word_bytes in {1,2,4,8}. */
int sparq_table_create(const void* host_words, uint64_t nwords, uint32_t word_bytes, sparq_table_t* out);
int sparq_table_destroy(sparq_table_t t);
int sparq_table_device_ptr(sparq_table_t t, const void** ptr);
/* Overwrite words [first, first+count) of a table from host memory (synchronous). */
int sparq_table_write(sparq_table_t t, uint64_t first, const void* host_words, uint64_t count);
/* Word 0 of a uint64 table += 1, ordered on the state's stream: the iteration
 * counter of SPARQ_PW_PAULI_TABLE.  Capturable (inside a captured sequence it
 * advances once per replay).  No counterpart in the reference. */
int sparq_table_advance(sparq_state_t st, sparq_table_t t);

/* ------------------------------------------------------------ operators */
/* Fused per-entry program (all key-rewrite and diagonal operators). */
int sparq_pointwise(sparq_state_t st, const sparq_pw_op_t* ops, uint32_t nops,
                    const sparq_cond_t* conds, uint32_t nconds);

/* ClearZero: stable removal of |amp|^2 < eps (system_operations.cpp:371-382). */
int sparq_clear_zero(sparq_state_t st, double eps);

/* ClearZero into caller-owned device arrays, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same result as sparq_clear_zero: the entries of keys_in /
 * amps_in ([words][cap_in] key words, word-major; n complex amplitudes as
 * re/im pairs, 16-byte aligned) with |amp|^2 >= eps, in order, go to keys_out /
 * amps_out ([words][cap_out]; cap_in, cap_out >= n), and their count to
 * *kept_dev (device memory).  Work is queued on `stream` (a hipStream_t; NULL
 * for the default stream) with no synchronization or readback.  Limit: n <=
 * 2^28 entries.  Per-tile counts and offsets live in a fixed ~1.5 MiB device
 * array of the library, one per GPU, shared by every call on that GPU: calls
 * that overlap on one GPU (different streams, or other threads) race. */
int sparq_clear_zero_noalloc(const uint64_t* keys_in, const double* amps_in, uint32_t words, uint64_t n, uint64_t cap_in,
                             double eps, uint64_t* keys_out, double* amps_out, uint64_t cap_out, uint64_t* kept_dev,
                             void* stream);

/* Sort by the full key, unsigned lexicographic over the words (SortUnconditional,
 * sort_state.cpp:84-88); when `last` is non-NULL the bits of that field are moved
 * to the least significant position (SortExceptKey, sort_state.cpp:10-36). */
int sparq_sort(sparq_state_t st, const sparq_field_t* last);

/* Stable sort by an ordering of fields: the `major` fields in the given order
 * (most significant first), then every key bit not in a listed field in
 * canonical order, then the `minor` fields in the given order.  Fields must not
 * overlap.  SortByKey (sort_state.cpp:46-53: major = {key}), SortByKey2
 * (:101-131: major = {key1, key2}), SortExceptKey (:10-36: minor = {key}),
 * SortExceptBit (:55-82: minor = {key without the bit, the bit}),
 * SortExceptKeyHadamard (:138-171: minor = the listed qubits of key, highest
 * first).  Entries equal under the ordering keep their order (the reference's
 * std::sort leaves it unspecified). */
int sparq_sort_by(sparq_state_t st, const sparq_field_t* major, uint32_t nmajor, const sparq_field_t* minor,
                  uint32_t nminor);
/* SortByAmplitude (sort_state.cpp:90-99): |amp|^2 descending, stable. */
int sparq_sort_by_amplitude(sparq_state_t st);
/* split_systems (system_operations.cpp:53-106): the entries of `st` failing
 * any condition move, in order, to `rest` (whose content is replaced); those
 * satisfying all stay in `st`, in order.  Both states have the same key width. */
int sparq_partition(sparq_state_t st, const sparq_cond_t* conds, uint32_t ncond, sparq_state_t rest);
/* combine_systems (system_operations.cpp:53-106): append the entries of `from`
 * (unchanged) after those of `to`. */
int sparq_append(sparq_state_t to, sparq_state_t from);

/* sort_merge_unique_erase: sort, sum amplitudes of equal keys, drop |sum|^2 < eps
 * (system_operations.h:157-181, basic_components.cpp:517-526). */
int sparq_merge_duplicates(sparq_state_t st, double eps);
/* add_systems(cur, add, coef) (system_operations.h:157-181). `add` is unchanged. */
int sparq_add_systems(sparq_state_t cur, sparq_state_t add, double coef, double eps);
/* CheckDuplicateKey (debugger.cpp:214-222): sorts; *has_dup = 1 when two entries share a key. */
int sparq_check_duplicates(sparq_state_t st, int* has_dup);

/* Hadamard_Int on the low n_digits bits of `target` (hadamard.cpp:18-162), followed by
 * ClearZero(eps).  Conditions are evaluated per entry. */
int sparq_hadamard_int(sparq_state_t st, sparq_field_t target, uint32_t n_digits,
                       const sparq_cond_t* conds, uint32_t ncond, double eps);
/* Hadamard_Int_Full (hadamard.cpp:181-338) incl. its CheckDuplicateKey, then ClearZero(eps). */
int sparq_hadamard_int_full(sparq_state_t st, sparq_field_t target,
                            const sparq_cond_t* conds, uint32_t ncond, double eps);

/* Rot_GeneralUnitary (SparQ/src/rot.cpp:340-467): every group of entries equal on all
 * registers but `target` is completed to 2^w entries and multiplied by the dense
 * 2^w x 2^w complex matrix `mat` (row-major, 2 doubles per element; w <= 6), then
 * ClearZero(eps).  Conditions are evaluated per entry. */
int sparq_group_unitary(sparq_state_t st, sparq_field_t target, const double* mat,
                        const sparq_cond_t* conds, uint32_t ncond, double eps);

/* Two-branch operators on bit `bit` of `target` (pairs of entries equal elsewhere). */
enum sparq_pair_kind {
    SPARQ_PAIR_HADAMARD_BOOL = 0,   /* Hadamard_Bool (hadamard.cpp:340-439) */
    SPARQ_PAIR_CONDROT_RATIONAL = 1,/* CondRot_Rational_Bool (condrot.cpp:12-151); angle from `in`, sign in mat[0].re */
    SPARQ_PAIR_CONDROT_TABLE = 2,   /* CondRot_General_Bool_Fast (condrot.h:149-332); 2x2 per value of `in` from table */
    SPARQ_PAIR_ROT_GENERAL = 3      /* Rot_Bool general branch (basic_gates.cpp:159-199) with mat */
};
typedef struct {
    uint32_t kind;
    sparq_field_t in;          /* angle / table index register */
    double mat[8];             /* 2x2 complex row-major for ROT_GENERAL; mat[0] = +1/-1 for CONDROT_RATIONAL */
    const double* table;       /* device: 8 doubles per value of `in` (CONDROT_TABLE) */
} sparq_pair_params_t;
int sparq_pair_op(sparq_state_t st, sparq_field_t target, uint32_t bit, const sparq_pair_params_t* params,
                  const sparq_cond_t* conds, uint32_t ncond, double eps);

/* AddRegisterWithHadamard (system_operations.cpp:264-290): every entry is copied
 * 2^width(target) times with the (zero) target field set to each value, amplitudes unchanged. */
int sparq_replicate(sparq_state_t st, sparq_field_t target);

/* Normalize (dark_magic.cpp:11-33). *norm2 receives the sum of |amp|^2 before scaling. */
int sparq_normalize(sparq_state_t st, double* norm2);
/* Sum of |amp|^2 without scaling (CheckNormalization / CheckNan, debugger.cpp:13-70: NaN iff an amplitude is NaN). */
int sparq_norm2(sparq_state_t st, double* norm2);
/* |amp|^2 summed over each tile of `tile` consecutive entries in entry order;
 * ntiles must be ceil(n / tile).  With sparq_state_read_range it samples an
 * entry with probability |amp|^2 (MeasureZ, SparQ/src/measurement.cpp:127-182). */
int sparq_tile_norm2(sparq_state_t st, uint64_t tile, double* out, uint64_t ntiles);
/* |amp|^2 summed over the entries satisfying every condition (Probability, measurement.cpp:281-300). */
int sparq_cond_norm2(sparq_state_t st, const sparq_cond_t* conds, uint32_t nconds, double* out);
/* Copy entries [first, first+count) to the host: keys entry-major (count*W words), amps re/im. */
int sparq_state_read_range(sparq_state_t st, uint64_t first, uint64_t count, uint64_t* keys, double* amps);
/* TestRemovable helper (debugger.cpp:191-212): *uniform = 1 when every entry has the same value. */
int sparq_field_uniform(sparq_state_t st, sparq_field_t f, int* uniform, uint64_t* value);

/* ------------------------------------------------------------ captured sequences
 * No counterpart in the reference, whose callers loop over operators with a
 * host synchronisation after every branching one.  A fixed operator sequence
 * is recorded once on a state and replayed as a HIP graph:
 *   sparq_capture_plan(st)     snapshot the state's keys; the caller then runs
 *                              the sequence once: nothing is launched, each
 *                              branching operator plans its transform from the
 *                              snapshot (segments: the distinct values of the
 *                              key bits that vary outside its target among the
 *                              entries its conditions select);
 *   sparq_capture_record(st)   allocate what the plans need, grow the state's
 *                              arrays to the capacity they imply and start
 *                              capturing the state's stream; the caller runs
 *                              the same sequence again;
 *   sparq_capture_end(st, &g)  stop capturing; g replays the sequence.
 * Inside a capture only sparq_pointwise (any program), sparq_hadamard_int /
 * sparq_hadamard_int_full (positive ClearZero threshold; conditions allowed),
 * sparq_table_advance, sparq_state_size / sparq_state_words and the capture
 * calls are accepted; anything else fails with SPARQ_INVALID (then call
 * sparq_capture_abort).  Recorded kernels read the entry count from device
 * memory and stride over it: replays need no host synchronisation.
 * sparq_graph_launch queues `times` replays on the state's stream; the next
 * call on the state reads the entry count and the violation flags, and fails
 * with SPARQ_RUNTIME if a replay broke an assumption of the plan (key bits
 * outside a branching operator's segments or constant bits, duplicate keys,
 * more entries than the capacity, a removed register not zero); the state's
 * content is then undefined.  A graph is bound to the state's arrays: after an
 * operation that reallocates them (an eager branching operator, a layout
 * change) sparq_graph_launch fails with SPARQ_INVALID.  Destroy a graph before
 * its state. */
typedef struct sparq_graph_s* sparq_graph_t;
int sparq_capture_plan(sparq_state_t st);
int sparq_capture_record(sparq_state_t st);
int sparq_capture_end(sparq_state_t st, sparq_graph_t* out);
int sparq_capture_abort(sparq_state_t st);
/* 0: no capture, 1: planning, 2: recording. */
int sparq_capture_phase(sparq_state_t st, int* phase);
int sparq_graph_launch(sparq_graph_t g, uint32_t times);
/* Entry capacity of the captured sequence and the graph's node count. */
int sparq_graph_info(sparq_graph_t g, uint64_t* capacity, uint64_t* nodes);
int sparq_graph_destroy(sparq_graph_t g);

/* ------------------------------------------------------------ diagnostics */
/* Message of the last failing call on this thread. */
const char* sparq_last_error(void);
/* Event timing helpers on a state's stream (used by the benchmark): record, then read ms. */
int sparq_timer_record(sparq_state_t st, int slot);
int sparq_timer_elapsed(sparq_state_t st, int slot_begin, int slot_end, float* ms);

/* Per-kernel profiling: when enabled, every engine kernel launch is bracketed by
 * HIP events on the stream it is launched on and credited with its algorithmic
 * HBM bytes.  sparq_prof_report synchronizes and returns one record per kernel
 * name (accumulated since the last reset). */
typedef struct {
    char name[48];
    uint64_t launches;
    double total_ms;
    double total_bytes;
} sparq_prof_record_t;
int sparq_prof_enable(int on);
int sparq_prof_reset(void);
int sparq_prof_report(sparq_prof_record_t* out, uint32_t capacity, uint32_t* count);

/* Device memory held by the engine's caching allocator (no counterpart in the
 * reference, whose CUDA path allocates through thrust): out[0] live bytes,
 * out[1] cached free bytes, out[2] peak live bytes, out[3] cache hits,
 * out[4] misses (runtime allocations), out[5] cached blocks.
 * sparq_release_cached_memory synchronizes the device and frees the cache. */
int sparq_memory_stats(uint64_t out[6]);
int sparq_release_cached_memory(void);

/* Post-selection into a dense vector, the device side of GetOutput
 * (SparQ_Algorithm/src/DiscreteAdiabatic/qda_fundamental.cpp:123-169): every
 * entry whose `zero` fields are all 0 and whose index (the `index` fields
 * concatenated with field 0 least significant) is below len is written to
 * out[index]; out holds len complex values (2*len doubles, re/im interleaved),
 * zero elsewhere, len <= 2^(sum of index widths) and <= 2^40 (GetOutput uses
 * the full 2^(sum of widths)); *weight = sum of |a|^2 of the selected entries
 * (the caller renormalises, as GetOutput does).  Keys are unique, so no index is
 * written twice. */
int sparq_post_select_dense(sparq_state_t st, const sparq_field_t* zero, uint32_t nzero, const sparq_field_t* index,
                            uint32_t nindex, double* out, uint64_t len, double* weight);

/* Post-selection in place, PartialTraceSelect / PartialTraceSelectRange
 * (SparQ/src/partial_trace.cpp:148-225): drop every entry that fails one of the
 * conditions (kept entries stay in order), then scale the kept amplitudes by
 * f = 1/sqrt(sum of their |a|^2).  *factor = f; +inf when nothing is kept
 * (the state is then empty), 0 for an empty input state. */
int sparq_select(sparq_state_t st, const sparq_cond_t* conds, uint32_t nconds, double* factor);

/* Radix-sort digit width (no counterpart in the reference, which sorts with
 * comparison sorts): 8 or 11 forces that width for every later sort, 0 (the
 * default) lets the engine's autotuner pick it per size class by timing the
 * first sorts of each class with both widths. */
int sparq_set_radix_bits(int bits);
/* Sort strategy (no counterpart in the reference): 0 (the default) moves whole
 * entries by the top 8-32 ordering bits in a few LSD passes and then sorts
 * chunks of <= 2048 entries in LDS (keys of <= 5 words, <= 128 ordering bits
 * below the prefix).  A bucket too large for one chunk makes the engine retry
 * with 8 more prefix bits (remembered per ordering and size class); anything
 * that still does not fit takes the LSD path.  1 always takes the LSD radix
 * passes over an index payload and a final gather.  Both give the same stable
 * order. */
int sparq_set_sort_mode(int mode);
/* Dense-domain Hadamard (no counterpart in the reference): with on != 0 (the
 * default) Hadamard_Int / Hadamard_Int_Full on a state whose low target bits
 * and other varying key bits span at most 2^26 slots, filled to at least a
 * quarter, run on a dense amplitude array (no grouping sort); 0 always takes
 * the grouped path.  Same states either way (amplitudes to rounding). */
int sparq_set_dense_transform(int on);
/* Output order of Hadamard_Int / Hadamard_Int_Full (no counterpart in the
 * reference): with on != 0 (the default), when every varying key bit outside
 * the transformed bits lies below them, the output is written in canonical
 * (target-major) order, so that a following sort finds it presorted; 0 keeps
 * the group-major order of the transform.  Same states either way. */
int sparq_set_canonical_output(int on);
/* Conditioned group unitaries (Rot_GeneralUnitary / Rot_GeneralStatePrep with
 * conditions; no counterpart in the reference): with on != 0 the groups whose
 * entries fail the conditions pass through inside the transform (the whole
 * state is grouped, the output stays in one group order); 0 (the default)
 * partitions the state first, groups and transforms only the satisfying
 * entries and appends the others behind them, which measured faster on the
 * QDA walk.  Same states either way. */
int sparq_set_group_pass_through(int on);
/* Entries per thread of the fused pointwise kernel: 1, 2 or 4 forces a shape
 * (4 only for keys of at most two words), 0 (default) lets the launch-shape
 * autotuner pick per class (key words, amplitude access, interpreter, log2 N)
 * from timed first calls.  Results do not depend on it. */
int sparq_set_pointwise_items(int items);
/* The autotuners' measurements and choices (pointwise launch shape, radix digit
 * width) as a JSON object in buf (NUL-terminated, truncated to len);
 * *needed receives the full length + 1. */
int sparq_autotune_report(char* buf, size_t len, size_t* needed);
/* Sort-path counters since the last reset (no counterpart in the reference):
 * out[0] sorts by the prefix + local path, out[1] their entries, out[2] local
 * attempts that overflowed a chunk, out[3] sorts by the LSD passes, out[4]
 * their entries, out[5] local sorts that needed no prefix pass (the state was
 * already sorted by a long enough key prefix).  count >= 6; reset != 0 clears
 * them after reading. */
int sparq_sort_stats(uint64_t* out, uint32_t count, int reset);

/* Learned output sizer (no counterpart in the reference, which always counts
 * before it allocates).  A 4-layer FP32 MLP, 16 features -> 64 -> 64 -> 64 -> 1
 * with ReLU, predicts log2(G)/32 for the group count G of Hadamard_Int,
 * Hadamard_Int_Full and dense group unitaries from the key-statistics pass.
 * Parameter layout (floats): W1[64][16], b1[64], W2[64][64], b2[64],
 * W3[64][64], b3[64], W4[64], b4.  When enabled, those operators size their
 * output with 2^(32*y + margin) groups and skip the readback of the exact group
 * count; the exact count, computed on the device, is read with the final size
 * and an under-prediction reruns the exact path (keys and amplitudes are
 * identical either way).  Training runs on the GPU: sparq_mlp_train performs
 * `steps` Adam steps on minibatches of a seeded permutation, loss
 * mean(w * (y_hat - y)^2) with w = alpha where y_hat < y, else 1. */
#define SPARQ_SIZER_FEATURES 16
#define SPARQ_SIZER_PARAMS 9473
int sparq_mlp_forward(const float* params, const float* x, uint32_t n, float* y);
int sparq_mlp_grad(const float* params, const float* x, const float* y, uint32_t n, float alpha, float* grad, float* loss);
int sparq_mlp_train(float* params, const float* x, const float* y, uint32_t n, uint32_t steps, uint32_t batch, float lr,
                    float alpha, uint64_t seed, float* loss_history);
int sparq_sizer_load(const float* params, uint32_t count, float margin);
int sparq_sizer_enable(int on);
/* Logging of (features, log2(G)/32) for every sized operator run on the exact path. */
int sparq_sizer_log(int on);
int sparq_sizer_log_fetch(float* x, float* y, uint32_t capacity, uint32_t* count, int clear);
/* out[0] operators sized by the exact count, out[1] by the network, out[2] network under-predictions (reruns). */
int sparq_sizer_stats(uint64_t out[3]);

#ifdef __cplusplus
}
#endif
#endif /* SPARQ_HIP_H */
