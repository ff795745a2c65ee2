/*
 * hip_batch_processing.h -- C++ drop-in for the reference's CuBatchProcessing
 * (cupro/CuBatchProcessing.cuh:47-112) as DevCommThread drives it
 * (cualn/DevCommThread.cpp:149-356), built on the C ABI of comer2_hip.h.
 *
 * The method names, argument lists and meaning follow the reference:
 *   TransferCPMData(bdbCpmbeg, bdbCpmend, szCpm2dvfields)   DB chunk upload
 *   SetDbDetails / SetQueryLen / SetScoreThreshold / SetLogEThreshold
 *   ProcessScoreMatrix(qrysernr, nqyposs, querypmbeg, bdb1pmbeg, bdb1pmend,
 *                      bdbCdesc, bdbCpmbeg, bdbCpmend, szCpm2dvfields, cnt)
 *   WaitForIdleChilds, GetCurrentMaxDbPos, GetCurrentMaxNDbPros and the
 *   Pass2 variants, GetCurrentDbxPadding(Pass2)
 * Field buffers are the reference's char*[pmv2DTotFlds] begin/end arrays
 * (cupro/PM2DVectorFields.h:41-81); they are borrowed and copied before each
 * call returns.  A DB made of a cached part (bdb1) and a chunk (bdbC) is
 * searched as one profile set, the bdb1 profiles first, as the reference
 * numbers them (cudp/CuBatchDP_init_btck.cu:96-110).
 *
 * Where the reference hands results to its finalizer thread and AlnWriter
 * (cupro/CuBatchProcessing.cu:596-694), this class calls the sink set with
 * SetResultsSink once per ProcessScoreMatrix, on the calling thread, with
 * the query serial number, the results (valid during the call) and bdbCdesc;
 * then it calls cnt->dec(), the release the reference's finalizer performs
 * after formatting (cupro/CuBatchProcessingFinalizer.cpp:186-190).  Counter
 * is any type with dec(), such as the reference's TSCounterVar.
 * Errors are thrown as comer2hip::Error, in place of myruntime_error.
 */
#ifndef HIP_BATCH_PROCESSING_H
#define HIP_BATCH_PROCESSING_H

#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include "comer2_hip.h"

namespace comer2hip {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }
private:
    int code_;
};

class HipBatchProcessing {
public:
    using Sink = std::function<void(int qrysernr, const cbp_results& res, const char** bdbCdesc)>;

    /* device, memory budget (0: the device's free memory), area number, the
     * options (src/env/options.txt, src/env/comer.par), the 30x30 SS table
     * and the 93 SS18 parameters (cupro/CuDeviceMemory.cu:40-80) */
    HipBatchProcessing(int device, size_t mem_budget, int areano, const cbp_options& opt,
                       const float* ssss900, const float* nnw93);
    ~HipBatchProcessing();
    HipBatchProcessing(const HipBatchProcessing&) = delete;
    HipBatchProcessing& operator=(const HipBatchProcessing&) = delete;

    void TransferCPMData(char** bdbCpmbeg, char** bdbCpmend, size_t* szCpm2dvfields);

    void SetDbDetails(uint64_t db_len, size_t no_sequences);
    size_t GetQueryLen() const { return querylength_; }
    void SetQueryLen(size_t value) { querylength_ = value; }
    float GetScoreThreshold() const { return scorethld_; }
    void SetScoreThreshold(float value) { scorethld_ = value; }
    float GetLogEThreshold() const { return logevthld_; }
    void SetLogEThreshold(float value) { logevthld_ = value; }

    template <class Counter>
    void ProcessScoreMatrix(int qrysernr, size_t nqyposs, char** querypmbeg,
                            char** bdb1pmbeg, char** bdb1pmend, const char** bdbCdesc,
                            char** bdbCpmbeg, char** bdbCpmend, size_t* szCpm2dvfields,
                            Counter* cnt)
    {
        Process(qrysernr, nqyposs, querypmbeg, bdb1pmbeg, bdb1pmend, bdbCdesc,
                bdbCpmbeg, bdbCpmend, szCpm2dvfields);
        if (cnt) cnt->dec();
    }

    /* results are delivered synchronously: nothing is left running */
    void WaitForIdleChilds() {}

    size_t GetCurrentMaxDbPos() const { return Limits(0); }
    size_t GetCurrentMaxNDbPros() const { return Limits(1); }
    size_t GetCurrentMaxDbPosPass2() const { return Limits(2); }
    size_t GetCurrentMaxNDbProsPass2() const { return Limits(3); }
    /* the score matrix is never materialised, so no column padding */
    unsigned int GetCurrentDbxPadding() const { return 0; }
    unsigned int GetCurrentDbxPaddingPass2() const { return 0; }

    void SetResultsSink(Sink sink) { sink_ = std::move(sink); }
    cbp_handle* handle() const { return h_; }

private:
    void Process(int qrysernr, size_t nqyposs, char** querypmbeg,
                 char** bdb1pmbeg, char** bdb1pmend, const char** bdbCdesc,
                 char** bdbCpmbeg, char** bdbCpmend, size_t* szCpm2dvfields);
    size_t Limits(int which) const;
    void Check(int rc, const char* what) const;

    cbp_handle* h_ = nullptr;
    size_t querylength_ = 0;
    float scorethld_ = 0.0f, logevthld_ = 0.0f;
    uint64_t dblen_ = 0;
    size_t nseqs_ = 0;
    /* the profile set on the device: (bdb1 begin, bdbC begin, bdbC end) of
     * the last upload */
    const void* up_bdb1_ = nullptr;
    const void* up_bdbC_ = nullptr;
    const void* up_bdbCend_ = nullptr;
    std::map<int, int32_t> qidx_;      /* query serial number -> query index */
    Sink sink_;
};

}  // namespace comer2hip

#endif
