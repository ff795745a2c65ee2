/*
 * c2_lib.h -- internal state behind the C ABI.
 */
#ifndef C2_LIB_H
#define C2_LIB_H
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <memory>
#include <map>
#include "../../include/comer2_hip.h"
#include "c2_prep.h"
#include "c2_kernels.h"

struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    ~DBuf() { if (p) (void)hipFree(p); }
    template <class T> T* as() const { return (T*)p; }
};

/* owned host copy of a profile set */
struct HostSet {
    cbp_profiles v{};
    std::vector<float> bkg, eno, trgfrq, ssprb, trans;
    std::vector<int32_t> len;
    std::vector<uint32_t> dist;
    std::vector<char> aa, sss;
    void copy_from(const cbp_profiles* s);
};

struct DevSet {
    int32_t nprof = 0;
    int64_t npos = 0;
    DBuf bkg, eno, len, dist, trgfrq, ssprb, aa, sss, trans, addrpro;
};

struct DevQuery {
    QueryPrep hp;
    DBuf qv, qtr, qtrb, QT, qh, lleft, ldiag, qbkg, qaa, mleft, mdiag;
};

/* per-query results of one cbp_process_batch call */
struct BatchStore {
    std::vector<float> aln, epa;
    std::vector<int32_t> order;
    std::vector<char> strs;
    std::vector<int> map;
};

struct cbp_handle {
    int device = 0;
    size_t budget = 0, used = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    cbp_options opt{};
    bool have_opt = false;
    std::vector<float> ssss;     /* 30x30 */
    std::vector<float> nnw;      /* 93 */
    bool have_tables = false;
    HostSet queries;
    bool have_queries = false;
    std::map<int, std::unique_ptr<DevQuery>> qcache;
    HostSet chunk;
    DevSet dchunk;
    bool have_chunk = false;
    std::vector<int32_t> seg_p;
    DBuf dseg_b, dseg_e;
    DBuf dnnw;           /* SS18 parameters on the device */
    /* per-call scratch */
    DBuf cand, cand_off, btck, fin, pass_pro, counts, statin, statout;
    DBuf mpairs, mF, mbt, mstat, malns, malntmp, mres;
    int32_t nseg = 0;
    DBuf keys;           /* [nprof] u64 */
    DBuf bot;            /* stripe hand-off */
    uint64_t dblen = 0, nseqs = 0;
    int seg_cols = 4096; /* tuning: target columns per segment */
    int rmax = 5;        /* tuning: max rows per lane */
    int nw = 8;          /* tuning: waves per workgroup */
    double last_ms = 0.0;
    int32_t last_nlaunch = 0;
    cbp_results res{};
    std::vector<float> res_aln;
    std::vector<int32_t> res_order;
    std::vector<char> res_strs;
    std::vector<float> res_epa;   /* log pair E-values before MAP reuses the slot */
    std::vector<int> res_map;     /* alndata rows realigned, in alns order */
    std::vector<BatchStore> batch; /* results of the last cbp_process_batch */
    /* soft DP: query set on the device and the prepared pair plan */
    DBuf sdq_trans, sdq_dist, sdq_ll, sdq_ld, sdq_trg, sdq_ss, sdq_bkg, dssss;
    bool sdq_ready = false;
    std::vector<SdpPair> sdp_pairs;
    std::vector<int32_t> sdp_order;
    DBuf sdp_dpairs, sdp_dorder;
    int32_t sdp_K = 0, sdp_nwaves = 0;
    int64_t sdp_ws_floats = 0, sdp_scr_off = 0;
    bool sdp_ready = false;
};

#endif
