/*
 * CPU oracle: layered offset/normalised min-sum LDPC decoder for the
 * DVB-S2/S2X IRA codes -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates lib/cusat_dsp/cuda/ldpc_decode_kernel.cu of wineslab/cusat:
 *   geometry / edge lists      :990-1065
 *   input clamp                :636-670 (int16 APP: in_lim = 32767)
 *   syndrome + early exit      :689-749
 *   read phase                 :770-781 (iteration 0 reads zero messages)
 *   check node (finalp_oms)    :242-275 and the helpers :33-58
 *   normalised scale hook      :75-90 (nms_scale, alpha in 64ths)
 *   message clamp              lib/cusat_dsp/cuda/ldpc_message_update.h:25-35
 *   data write (re-reads APP)  :793-801
 *   parity write (read value)  :803-833
 *   hard-decision pack         :468-493
 *
 * Serialisation.  The reference kernel updates shared-memory APPs from 720
 * threads with no barrier between a step's reads and writes; when one
 * variable node is touched twice in a step (a weight-2/3 circulant inside a
 * layer, or a variable-node group shared by the two layers of a step) its
 * outcome depends on thread timing.  This oracle fixes one legal execution:
 * every read of a step (both lanes) sees the APP state at the start of the
 * step; then the data writes are applied lane by lane, and within a lane
 * circulant by circulant in table order (each circulant writes 360 distinct
 * variable nodes), every write re-reading the current APP with int16
 * saturation per add; then the parity writes.  Wherever no variable node
 * is touched twice in a step this is exactly the reference's arithmetic.
 * The edge order inside one check is the circulant order; the check node
 * result does not depend on edge order (min1/min2/sign-XOR are symmetric).
 */
#include "oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define M 360
#define MAX_DEG 64

struct oc_code {
    int N, K, R, q, n_rows;
    int links_total, links_max_cn;
    int* row_deg;
    int* row_off;
    int* addr;
    /* per layer: circulants in table order (row r asc, address index asc) */
    int* lay_cnt;   /* [q] */
    int* lay_off;   /* [q+1] */
    int* circ_row;  /* variable-node group r (VN = 360 r + jj) */
    int* circ_shift;/* s = a div q: check j connects to jj = (j - s) mod 360 */
    int n_steps;
    int (*steps)[2];
};

/* ---- table text parsing (format written by
 *      cusat_amd/data/extract_reference_data.py) ---- */

static char* slurp(const char* path)
{
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char* b = (char*)malloc(n + 1);
    if (fread(b, 1, n, f) != (size_t)n) { fclose(f); free(b); return NULL; }
    b[n] = 0;
    fclose(f);
    return b;
}

static int load_table(oc_code* c, const char* text, const char* name)
{
    char hdr[64];
    snprintf(hdr, sizeof hdr, "@table\n%s N=", name);
    const char* p = strstr(text, hdr);
    if (!p) return -1;
    p += strlen("@table\n");
    char nm[32];
    int mm;
    if (sscanf(p, "%31s N=%d K=%d M=%d", nm, &c->N, &c->K, &mm) != 4 || mm != M)
        return -1;
    p = strchr(p, '\n') + 1;
    c->R = c->N - c->K;
    c->q = c->R / M;
    c->n_rows = c->K / M;
    c->row_deg = (int*)calloc(c->n_rows, sizeof(int));
    c->row_off = (int*)calloc(c->n_rows + 1, sizeof(int));
    c->addr = (int*)malloc(sizeof(int) * 32 * c->n_rows);
    int na = 0;
    for (int r = 0; r < c->n_rows; r++) {
        const char* e = strchr(p, '\n');
        c->row_off[r] = na;
        while (p < e) {
            char* q2;
            long v = strtol(p, &q2, 10);
            if (q2 == p) break;
            c->addr[na++] = (int)v;
            p = q2;
            while (*p == ' ') p++;
        }
        c->row_deg[r] = na - c->row_off[r];
        p = e + 1;
    }
    c->row_off[c->n_rows] = na;
    if (sscanf(p, "@links %d %d", &c->links_total, &c->links_max_cn) != 2) return -1;
    return 0;
}

static void load_sched(oc_code* c, const char* text, const char* name)
{
    char hdr[64];
    snprintf(hdr, sizeof hdr, "@sched %s ", name);
    const char* p = text ? strstr(text, hdr) : NULL;
    int q = 0, n = 0;
    if (p) {
        const char* qq = strstr(p, "q=");
        sscanf(qq, "q=%d steps=%d", &q, &n);
    }
    if (p && q == c->q) {
        c->n_steps = n;
        c->steps = malloc(sizeof(int[2]) * n);
        p = strchr(p, '\n') + 1;
        for (int s = 0; s < n; s++) {
            sscanf(p, "%d %d", &c->steps[s][0], &c->steps[s][1]);
            p = strchr(p, '\n') + 1;
        }
    } else {
        /* sequential fallback: one layer per step, lane 1 idle
         * (ldpc_schedule_registry.h:102-106) */
        c->n_steps = c->q;
        c->steps = malloc(sizeof(int[2]) * c->q);
        for (int i = 0; i < c->q; i++) { c->steps[i][0] = i; c->steps[i][1] = -1; }
    }
}

oc_code* oc_code_open(const char* tables_txt, const char* sched_txt,
                      const char* name, int use_sched)
{
    char* tt = slurp(tables_txt);
    if (!tt) return NULL;
    oc_code* c = (oc_code*)calloc(1, sizeof(oc_code));
    if (load_table(c, tt, name)) { free(tt); oc_code_close(c); return NULL; }
    free(tt);
    char* st = (use_sched && sched_txt) ? slurp(sched_txt) : NULL;
    load_sched(c, st, name);
    free(st);

    /* Layer i holds the checks c = q*j + i; address a of row r feeds check
     * (a + jj*q) mod R for bit 360 r + jj (ldpc_decode_kernel.cu:1006-1019),
     * so a belongs to layer a mod q with shift a div q. */
    c->lay_cnt = (int*)calloc(c->q, sizeof(int));
    c->lay_off = (int*)calloc(c->q + 1, sizeof(int));
    int total = c->row_off[c->n_rows];
    for (int n = 0; n < total; n++) c->lay_cnt[c->addr[n] % c->q]++;
    for (int i = 0; i < c->q; i++) c->lay_off[i + 1] = c->lay_off[i] + c->lay_cnt[i];
    c->circ_row = (int*)malloc(sizeof(int) * total);
    c->circ_shift = (int*)malloc(sizeof(int) * total);
    int* fill = (int*)calloc(c->q, sizeof(int));
    for (int r = 0; r < c->n_rows; r++)
        for (int n = c->row_off[r]; n < c->row_off[r + 1]; n++) {
            int a = c->addr[n], i = a % c->q;
            int k = c->lay_off[i] + fill[i]++;
            c->circ_row[k] = r;
            c->circ_shift[k] = a / c->q;
        }
    free(fill);
    return c;
}

void oc_code_close(oc_code* c)
{
    if (!c) return;
    free(c->row_deg); free(c->row_off); free(c->addr);
    free(c->lay_cnt); free(c->lay_off); free(c->circ_row); free(c->circ_shift);
    free(c->steps);
    free(c);
}

void oc_code_info(const oc_code* c, int* N, int* K, int* q, int* n_steps,
                  int* cnl, int* lt_alloc)
{
    if (N) *N = c->N;
    if (K) *K = c->K;
    if (q) *q = c->q;
    if (n_steps) *n_steps = c->n_steps;
    if (cnl) *cnl = c->links_max_cn - 2;
    if (lt_alloc) {
        /* bnl_layer_off[q] = sum_i M (cnt_i + 2) (ldpc_decode_kernel.cu:1052-1065) */
        int s = 0;
        for (int i = 0; i < c->q; i++) s += M * (c->lay_cnt[i] + 2);
        *lt_alloc = s;
    }
}

/* ---- fixed-point helpers (ldpc_decode_kernel.cu:33-58) ---- */

static inline int16_t qabs(int16_t a) { return a == -32768 ? 32767 : (a < 0 ? -a : a); }
static inline uint16_t qsubu(uint16_t a, uint16_t b) { return a > b ? a - b : 0; }
static inline int16_t qsat16(int v)
{
    return (int16_t)(v < -32768 ? -32768 : (v > 32767 ? 32767 : v));
}

int oc_clamp_msg(int v) { return v < -127 ? -127 : (v > 127 ? 127 : v); }

/* nms_scale (ldpc_decode_kernel.cu:75-90): odd-symmetric, alpha = n/64 */
static inline int nms_scale(int v, int n)
{
    if (n == 0) return v;
    const int m = ((v < 0 ? -v : v) * n + 32) >> 6;
    return v < 0 ? -m : m;
}

/* finalp_oms_i16 (ldpc_decode_kernel.cu:242-275) followed by nms_scale and
 * the int8 clamp applied before every store (:796,806,817). */
void oc_check_node(const int16_t* x, int deg, int beta, int alpha64, int16_t* out)
{
    const uint16_t b = (uint16_t)beta;
    int16_t min1, min2, xs = x[0];
    {
        uint16_t m0 = qsubu((uint16_t)qabs(x[0]), b);
        uint16_t m1 = qsubu((uint16_t)qabs(x[1]), b);
        xs ^= x[1];
        if (m0 <= m1) { min1 = m0; min2 = m1; }
        else          { min1 = m1; min2 = m0; }
    }
    for (int e = 2; e < deg; e++) {
        xs ^= x[e];
        int16_t m = (int16_t)qsubu((uint16_t)qabs(x[e]), b);
        if (m < min1) { min2 = min1; min1 = m; }
        else if (m < min2) { min2 = m; }
    }
    for (int e = 0; e < deg; e++) {
        int16_t m = (int16_t)qsubu((uint16_t)qabs(x[e]), b);
        int16_t oth = (m == min1) ? min2 : min1;
        int neg = ((int16_t)(xs ^ x[e])) < 0;
        int o = neg ? -(int)oth : (int)oth;
        out[e] = (int16_t)oc_clamp_msg(nms_scale(o, alpha64));
    }
}

/* ---- decoder ---- */

typedef struct {
    int16_t* app_d;  /* [K] */
    int16_t* app_p;  /* [R], parity bit c <-> channel index K + c */
    int8_t* msg;     /* [R][MAX_DEG]: edges k < cnt data, cnt = own parity,
                        cnt+1 = previous parity */
    int16_t* x;      /* per step: [2][M][MAX_DEG] read-phase values */
    int16_t* y;      /* per step: [2][M][MAX_DEG] new messages */
} work_t;

static inline int check_deg(const oc_code* c, int cidx, int i)
{
    return c->lay_cnt[i] + (cidx == 0 ? 1 : 2);
}

static inline int vn_of(const oc_code* c, int k, int j)
{
    return M * c->circ_row[k] + ((j - c->circ_shift[k] + M) % M);
}

static int syndrome_ok(const oc_code* c, const int16_t* app_d, const int16_t* app_p)
{
    for (int i = 0; i < c->q; i++)
        for (int j = 0; j < M; j++) {
            int cidx = c->q * j + i;
            int16_t syn = app_p[cidx];
            if (cidx > 0) syn ^= app_p[cidx - 1];
            for (int k = c->lay_off[i]; k < c->lay_off[i + 1]; k++)
                syn ^= app_d[vn_of(c, k, j)];
            if (syn < 0) return 0;
        }
    return 1;
}

int oc_ldpc_syndrome_ok(const oc_code* c, const int16_t* llr)
{
    return syndrome_ok(c, llr, llr + c->K);
}

static void decode_one(const oc_code* c, work_t* w, const int16_t* llr,
                       int T, const int16_t* beta, const int16_t* alpha,
                       int n_param_iters, uint8_t* hard, int32_t* result)
{
    const int K = c->K, R = c->R, q = c->q;
    for (int v = 0; v < K; v++) w->app_d[v] = llr[v] < -32767 ? -32767 : llr[v];
    for (int p = 0; p < R; p++) w->app_p[p] = llr[K + p] < -32767 ? -32767 : llr[K + p];
    memset(w->msg, 0, (size_t)R * MAX_DEG);

    int res = -1;
    for (int t = 0; t <= T; t++) {
        if (syndrome_ok(c, w->app_d, w->app_p)) { res = T - t; break; }
        if (t == T) break;
        for (int s = 0; s < c->n_steps; s++) {
            /* read phase + check node, both lanes, on the step-start state */
            for (int lane = 0; lane < 2; lane++) {
                int i = c->steps[s][lane];
                if (i < 0) continue;
                int bt = 1, al = 0;
                if (beta && t < n_param_iters) bt = beta[t * q + i];
                if (alpha && t < n_param_iters) al = alpha[t * q + i];
                int cnt = c->lay_cnt[i];
                for (int j = 0; j < M; j++) {
                    int cidx = q * j + i;
                    int deg = check_deg(c, cidx, i);
                    int8_t* mo = w->msg + (size_t)cidx * MAX_DEG;
                    int16_t* x = w->x + ((size_t)lane * M + j) * MAX_DEG;
                    int16_t* y = w->y + ((size_t)lane * M + j) * MAX_DEG;
                    for (int e = 0; e < cnt; e++)
                        x[e] = qsat16((int)w->app_d[vn_of(c, c->lay_off[i] + e, j)] - mo[e]);
                    x[cnt] = qsat16((int)w->app_p[cidx] - mo[cnt]);
                    if (deg > cnt + 1)
                        x[cnt + 1] = qsat16((int)w->app_p[cidx - 1] - mo[cnt + 1]);
                    oc_check_node(x, deg, bt, al, y);
                }
            }
            /* data writes: lane by lane, circulant by circulant */
            for (int lane = 0; lane < 2; lane++) {
                int i = c->steps[s][lane];
                if (i < 0) continue;
                int cnt = c->lay_cnt[i];
                for (int e = 0; e < cnt; e++)
                    for (int j = 0; j < M; j++) {
                        int cidx = q * j + i;
                        int v = vn_of(c, c->lay_off[i] + e, j);
                        int8_t old = w->msg[(size_t)cidx * MAX_DEG + e];
                        int16_t yy = w->y[((size_t)lane * M + j) * MAX_DEG + e];
                        w->app_d[v] = qsat16((int)w->app_d[v] + yy - old);
                    }
            }
            /* parity writes from the read-phase values, then store messages */
            for (int lane = 0; lane < 2; lane++) {
                int i = c->steps[s][lane];
                if (i < 0) continue;
                int cnt = c->lay_cnt[i];
                for (int j = 0; j < M; j++) {
                    int cidx = q * j + i;
                    int deg = check_deg(c, cidx, i);
                    const int16_t* x = w->x + ((size_t)lane * M + j) * MAX_DEG;
                    const int16_t* y = w->y + ((size_t)lane * M + j) * MAX_DEG;
                    w->app_p[cidx] = qsat16((int)x[cnt] + y[cnt]);
                    if (deg > cnt + 1)
                        w->app_p[cidx - 1] = qsat16((int)x[cnt + 1] + y[cnt + 1]);
                    int8_t* mo = w->msg + (size_t)cidx * MAX_DEG;
                    for (int e = 0; e < deg; e++) mo[e] = (int8_t)y[e];
                }
            }
        }
    }
    memset(hard, 0, K / 8);
    for (int v = 0; v < K; v++)
        if (w->app_d[v] < 0) hard[v >> 3] |= (uint8_t)(0x80 >> (v & 7));
    *result = res;
}

static work_t* work_new(const oc_code* c)
{
    work_t* w = (work_t*)malloc(sizeof(work_t));
    w->app_d = (int16_t*)malloc(sizeof(int16_t) * c->K);
    w->app_p = (int16_t*)malloc(sizeof(int16_t) * c->R);
    w->msg = (int8_t*)malloc((size_t)c->R * MAX_DEG);
    w->x = (int16_t*)malloc(sizeof(int16_t) * 2 * M * MAX_DEG);
    w->y = (int16_t*)malloc(sizeof(int16_t) * 2 * M * MAX_DEG);
    return w;
}

static void work_free(work_t* w)
{
    free(w->app_d); free(w->app_p); free(w->msg); free(w->x); free(w->y); free(w);
}

int oc_ldpc_decode(const oc_code* c, const int16_t* llr, int max_trials,
                   const int16_t* beta, const int16_t* alpha, int n_param_iters,
                   uint8_t* hard, int32_t* result)
{
    if (!c || !llr || !hard || !result || max_trials < 0) return -1;
    if (c->links_max_cn > MAX_DEG) return -1;
    work_t* w = work_new(c);
    decode_one(c, w, llr, max_trials, beta, alpha, n_param_iters, hard, result);
    work_free(w);
    return 0;
}

typedef struct {
    const oc_code* c;
    const int16_t* llr;
    int B, T, n_param_iters, nthreads, tid;
    const int16_t *beta, *alpha;
    uint8_t* hard;
    int32_t* result;
} job_t;

static void* worker(void* arg)
{
    job_t* jb = (job_t*)arg;
    work_t* w = work_new(jb->c);
    for (int b = jb->tid; b < jb->B; b += jb->nthreads)
        decode_one(jb->c, w, jb->llr + (size_t)b * jb->c->N, jb->T, jb->beta,
                   jb->alpha, jb->n_param_iters, jb->hard + (size_t)b * (jb->c->K / 8),
                   jb->result + b);
    work_free(w);
    return NULL;
}

int oc_ldpc_decode_batch(const oc_code* c, const int16_t* llr, int B,
                         int max_trials, const int16_t* beta,
                         const int16_t* alpha, int n_param_iters,
                         uint8_t* hard, int32_t* result, int nthreads)
{
    if (!c || B < 0 || max_trials < 0) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > B) nthreads = B > 0 ? B : 1;
    pthread_t th[256];
    job_t jobs[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (job_t){ c, llr, B, max_trials, n_param_iters, nthreads, t,
                           beta, alpha, hard, result };
        pthread_create(&th[t], NULL, worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    return 0;
}
