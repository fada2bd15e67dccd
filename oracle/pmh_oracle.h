/*
 * pmh_oracle.h -- CPU oracle interface (test infrastructure only; see
 * pmh_oracle.c).  Tag and opcode numbering matches include/pmh.h so the
 * tests can feed the same descriptors to both.
 */
#ifndef PMH_ORACLE_H
#define PMH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_I64 = 0, ORC_F64 = 1, ORC_BOOL = 2, ORC_I32 = 3, ORC_F32 = 4, ORC_I8 = 5,
    ORC_I16 = 6, ORC_U8 = 7, ORC_U16 = 8, ORC_U32 = 9, ORC_U64 = 10
};
enum { ORC_EQ = 0, ORC_NE = 1, ORC_LT = 2, ORC_LE = 3, ORC_GT = 4, ORC_GE = 5 };
enum {
    ORC_N_COLUMN = 0, ORC_N_LIT_I64 = 1, ORC_N_LIT_F64 = 2, ORC_N_LIT_BOOL = 3,
    ORC_N_COMPARE = 4, ORC_N_AND = 5, ORC_N_OR = 6
};
#define ORC_MAX_VALS 16

typedef struct {
    int32_t tag;
    int32_t pad;
    const void *data;
    const uint8_t *valid; /* NULL = all valid */
} orc_col;

/* One postfix predicate node.  dtype is the operand type for COLUMN and
 * COMPARE nodes (ORC_I64, ORC_F64 or ORC_BOOL). */
typedef struct {
    int32_t kind;
    int32_t op;
    int32_t col;
    int32_t dtype;
    int64_t ival;
    double fval;
} orc_node;

typedef struct orc_gb orc_gb;

int orc_data_bits(int tag);
int orc_eval_predicate(const orc_node *nodes, int n_nodes, const orc_col *cols,
                       uint64_t n_rows, uint8_t *out_data, uint8_t *out_valid);
uint64_t orc_compact(const uint8_t *keep, uint64_t n_rows, int elem_size,
                     const void *data, const uint8_t *valid,
                     void *out_data, uint8_t *out_valid);
double orc_exact_sum(const double *x, uint64_t n);
int orc_encode_keys(int n_keys, const orc_col *keys, uint64_t n_rows,
                    uint64_t *lo, uint64_t *hi);
orc_gb *orc_groupby(uint64_t n_rows, const uint8_t *keep, int n_keys, const orc_col *keys,
                    int n_vals, const orc_col *vals, int sum_mode);
orc_gb *orc_groupby_par(uint64_t n_rows, const uint8_t *keep, int n_keys, const orc_col *keys,
                        int n_vals, const orc_col *vals, int sum_mode, int n_threads);
uint64_t orc_gb_n_groups(const orc_gb *g);
void orc_gb_lanes(const orc_gb *g, uint64_t *lo, uint64_t *hi, uint32_t *first_row, uint32_t *len);
void orc_gb_val(const orc_gb *g, int j, uint32_t *count, int64_t *sum_i, int64_t *min_i,
                int64_t *max_i, double *sum_f, double *min_f, double *max_f);
void orc_gb_free(orc_gb *g);

#ifdef __cplusplus
}
#endif
#endif
