/*
 * c2_batchproc.cpp -- comer2hip::HipBatchProcessing, the C++ drop-in for the
 * reference's CuBatchProcessing (cupro/CuBatchProcessing.cuh:47-112) over
 * the C ABI (include/hip_batch_processing.h).
 */
#include <string>
#include "../../include/hip_batch_processing.h"

namespace comer2hip {

void HipBatchProcessing::Check(int rc, const char* what) const
{
    if (rc == CBP_OK) return;
    std::string msg = std::string("HipBatchProcessing::") + what + ": ";
    msg += h_ ? cbp_last_error(h_) : "no handle";
    throw Error(rc, msg);
}

HipBatchProcessing::HipBatchProcessing(int device, size_t mem_budget, int areano, const cbp_options& opt,
                                       const float* ssss900, const float* nnw93)
{
    const int rc = cbp_create(device, mem_budget, areano, &h_);
    if (rc != CBP_OK) throw Error(rc, "HipBatchProcessing: cbp_create failed (no usable GPU or HIP error)");
    try {
        Check(cbp_set_options(h_, &opt), "options");
        Check(cbp_upload_tables(h_, ssss900, nnw93, nnw93 ? 93 : 0), "tables");
    } catch (...) {
        cbp_destroy(h_);
        h_ = nullptr;
        throw;
    }
}

HipBatchProcessing::~HipBatchProcessing()
{
    if (h_) cbp_destroy(h_);
}

/* DevCommThread::ProcessScoreMatrix passes the chunk again with every query;
 * DevCommThread uploads it once per chunk through TransferCPMData
 * (cualn/DevCommThread.cpp:296-307) */
void HipBatchProcessing::TransferCPMData(char** bdbCpmbeg, char** bdbCpmend, size_t* szCpm2dvfields)
{
    (void)szCpm2dvfields;      /* field extents come from the end pointers */
    if (!bdbCpmbeg || !bdbCpmend) throw Error(CBP_ERR_ARG, "HipBatchProcessing::TransferCPMData: null data");
    Check(cbp_upload_chunk_fields(h_, (const void* const*)bdbCpmbeg, (const void* const*)bdbCpmend),
          "TransferCPMData");
    up_bdb1_ = nullptr;
    up_bdbC_ = bdbCpmbeg[0];
    up_bdbCend_ = bdbCpmend[0];
}

void HipBatchProcessing::SetDbDetails(uint64_t db_len, size_t no_sequences)
{
    dblen_ = db_len;
    nseqs_ = no_sequences;
    Check(cbp_set_db_details(h_, db_len, (uint64_t)no_sequences), "SetDbDetails");
}

size_t HipBatchProcessing::Limits(int which) const
{
    uint64_t v[4] = {0, 0, 0, 0};
    const uint64_t qlen = querylength_ ? querylength_ : 1;
    Check(cbp_chunk_limits(h_, qlen, &v[0], &v[1], &v[2], &v[3]), "chunk limits");
    return (size_t)v[which];
}

void HipBatchProcessing::Process(int qrysernr, size_t nqyposs, char** querypmbeg,
                                 char** bdb1pmbeg, char** bdb1pmend, const char** bdbCdesc,
                                 char** bdbCpmbeg, char** bdbCpmend, size_t* szCpm2dvfields)
{
    (void)szCpm2dvfields;
    if (!querypmbeg) throw Error(CBP_ERR_ARG, "HipBatchProcessing::ProcessScoreMatrix: null query data");
    if (!bdb1pmbeg && !bdbCpmbeg) throw Error(CBP_ERR_ARG, "HipBatchProcessing::ProcessScoreMatrix: no DB data");
    /* the profile set on the device must be bdb1 followed by bdbC */
    const void* b1 = bdb1pmbeg ? bdb1pmbeg[0] : nullptr;
    const void* bC = bdbCpmbeg ? bdbCpmbeg[0] : nullptr;
    const void* eC = bdbCpmend ? bdbCpmend[0] : nullptr;
    if (b1 != up_bdb1_ || bC != up_bdbC_ || eC != up_bdbCend_) {
        Check(cbp_upload_db_fields(h_, (const void* const*)bdb1pmbeg, (const void* const*)bdb1pmend,
                                   (const void* const*)bdbCpmbeg, (const void* const*)bdbCpmend),
              "ProcessScoreMatrix (DB upload)");
        up_bdb1_ = b1; up_bdbC_ = bC; up_bdbCend_ = eC;
    }
    auto it = qidx_.find(qrysernr);
    int32_t qi;
    if (it == qidx_.end()) {
        Check(cbp_add_query_fields(h_, (const void* const*)querypmbeg, (uint64_t)nqyposs, &qi),
              "ProcessScoreMatrix (query)");
        qidx_[qrysernr] = qi;
    } else {
        qi = it->second;
    }
    cbp_results res;
    Check(cbp_process(h_, qi, scorethld_, logevthld_, &res), "ProcessScoreMatrix");
    if (sink_) sink_(qrysernr, res, bdbCdesc);
}

}  // namespace comer2hip
