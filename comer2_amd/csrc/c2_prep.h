/*
 * c2_prep.h -- host-side per-query preparation for the phase-1 kernel.
 */
#ifndef C2_PREP_H
#define C2_PREP_H
#include <vector>
#include <stdint.h>
#include "../../include/comer2_hip.h"

struct QueryPrep {
    int Lq = 0, R = 1, nstripes = 1, Lpad = 0;
    std::vector<float> qv, qtr, qtrb, QT, qh, lleft, ldiag, qbkg;
    std::vector<float> mleft, mdiag;   /* [5][Lq] MAP forward column -1 boundary */
};

/* stripe geometry for a query of length Lq with at most rmax rows per lane */
void c2_geometry(int Lq, int rmax, int* R, int* nstripes);

/* fill a QueryPrep for profile qp of set q; ssss = 30x30 table */
void c2_prepare_query(const cbp_profiles* q, int qp, const float* ssss, int rmax, QueryPrep* out);

/* MAP forward column -1 boundary (see c2_prep.cpp) */
void c2_map_colm1(const cbp_profiles* q, int qp, float* mleft /*[Lq][5]*/, float* mdiag /*[Lq][5]*/);

/* column -1 boundary (see c2_prep.cpp) */
void c2_colm1(const cbp_profiles* q, int qp, float* lleft /*[Lq][5]*/, float* ldiag /*[Lq][5]*/);

#endif
