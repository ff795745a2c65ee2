/*
 * comer2_hip.h -- C ABI of the MI355X (gfx950) libmycu search path.
 *
 * The entry points replace the device side of the reference's
 * CuBatchProcessing (cupro/CuBatchProcessing.cuh:47-244) as driven by
 * DevCommThread (cualn/DevCommThread.cpp:149-356).  Plain C: opaque handle,
 * borrowed host pointers (copied before return), int status codes
 * (0 = success), no exceptions across the boundary, one handle per worker
 * thread (calls on one handle are not thread-safe; distinct handles are).
 */
#ifndef COMER2_HIP_H
#define COMER2_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CBP_OK            0
#define CBP_ERR_ARG      -1   /* invalid argument */
#define CBP_ERR_HIP      -2   /* HIP runtime failure */
#define CBP_ERR_MEM      -3   /* device memory budget exceeded */
#define CBP_ERR_STATE    -4   /* call out of order (e.g. no query uploaded) */
#define CBP_ERR_NODEV    -5   /* no usable GPU */

/* number of alignment-data fields per passed pair; field order of
 * TDP2OutputAlnData, cuss/CuBatchSS_com.h:45-71 */
enum {
    CBP_F_SCORE = 0, CBP_F_ORGPRONO, CBP_F_PRONEWDST, CBP_F_PSTVS,
    CBP_F_BEGCOORDS, CBP_F_ENDCOORDS,
    CBP_F_LMBDEST, CBP_F_KEST, CBP_F_EVALUE, CBP_F_EPA,
    CBP_F_MIN, CBP_F_MAX, CBP_F_E, CBP_F_H, CBP_F_LMBD, CBP_F_K,
    CBP_F_IDNTS, CBP_F_NGAPS,
    CBP_NFIELDS
};

/* SoA profile set; mirrors the fields of PM2DVectorFields
 * (cupro/PM2DVectorFields.h:48-81).  Field-major arrays:
 *   bkg    [20][nprof]      background probabilities (pps2DBkgPrbs)
 *   eno    [nprof]          effective number of observations (pps2DENO)
 *   len    [nprof]          profile length (pps2DLen)
 *   dist   [nprof]          first position of each profile (pps2DDist)
 *   trgfrq [20][npos]       target frequencies (pmv2DTrgFrqs)
 *   ssprb  [3][npos]        SS state probabilities (pmv2DSSsprbs)
 *   aa     [npos]           residue letters (pmv2Daa)
 *   sss    [npos]           SS state (pmv2DSSstate)
 *   trans  [7][npos+nprof]  log transitions MM MI MD IM II DM DD, len+1 rows
 *                           per profile, row 0 = beginning (ptr2DTrnPrbs)
 *   cv     [19][npos]       context vectors (pmv2DCVentrs)
 *   cvprior [npos]          log prior of the context vector (pmv2DCVprior)
 *   cvnorm2 [npos]          squared norm of the context vector (pmv2DCVnorm2)
 * cv, cvprior and cvnorm2 are NULL together when the profiles carry no
 * context vectors; the CV score term then sees the reader's uninformative
 * vector (every entry 0.17302947, prior 0, norm 0;
 * cupro/IOProfileModel.cu:700-707) */
typedef struct {
    int32_t         nprof;
    int64_t         npos;
    const float*    bkg;
    const float*    eno;
    const int32_t*  len;
    const uint32_t* dist;
    const float*    trgfrq;
    const float*    ssprb;
    const char*     aa;
    const char*     sss;
    const float*    trans;
    const float*    cv;
    const float*    cvprior;
    const float*    cvnorm2;
} cbp_profiles;

/* number of context-vector entries per position (pmv2DNoCVEls) */
#define CBP_NCVEL 19

/* options (src/env/options.txt) and statistical parameters (src/env/comer.par) */
typedef struct {
    float   ssswgt;     /* SSSWGT  options.txt:81 */
    int32_t mapaln;     /* MAPALN  options.txt:126 */
    float   minpp;      /* MINPP   options.txt:131 */
    float   evalue;     /* EVAL    options.txt:9 */
    int32_t ssemodel;   /* SSEMODEL options.txt:93 */
    int32_t xuninf;     /* X_UNINF options.txt:28 */
    float   ung_lambda, ung_K, ung_H;             /* [UNGAPPED]  comer.par:5-11 */
    float   gap_lambda, gap_K, gap_alpha, gap_beta; /* [GAPPED_A_1] comer.par:15-20 */
    int32_t max_map_pairs; /* cap on realigned pairs per query (<=0: no cap) */
    float   cvswgt;     /* CVSWGT  options.txt:74; > 0 adds the context-vector
                           score term (needs cbp_upload_cvs2s) */
} cbp_options;

/* per-query search outcome; buffers are owned by the handle and stay valid
 * until the next cbp_process* call on it */
typedef struct {
    int32_t      npassed;    /* pairs passing phase 1 (DP + correlation) */
    int64_t      nposits;    /* sum of passed profile lengths */
    int32_t      maxlen;     /* longest passed profile */
    int32_t      nmap;       /* pairs drawn: realigned by MAP (MAPALN=1) or drawn
                                from the phase-1 backtrace (MAPALN=0), the
                                selected pairs in order */
    const float* alndata;    /* [npassed][CBP_NFIELDS] in profile order */
    const int32_t* order;    /* [npassed] indices into alndata sorted by (log E, profile index) */
    const char*  alns;       /* alignment strings: per realigned pair, 3 rows
                                of alnstride chars (query, middle, target),
                                NUL-terminated */
    int64_t      alnstride;
    const int32_t* map_rows; /* [nmap] alndata row of each realigned pair */
    const float* epa;        /* [npassed] log pair E-value from the statistics
                                (the EPA field holds the MAP alignment length,
                                as in the reference's output records) */
    float        scorethld, logevthld;
    double       search_space;
} cbp_results;

typedef struct cbp_handle cbp_handle;

/* lifecycle (replaces CuDeviceMemory construction, cupro/CuDeviceMemory.cu:85-146) */
int  cbp_create(int device, size_t mem_budget, int area, cbp_handle** out);
int  cbp_destroy(cbp_handle* h);
const char* cbp_last_error(const cbp_handle* h);

/* options + tables (CuDeviceMemory::CacheCompleteData and its CacheSSENNWeights /
 * CacheSSSScores, cupro/CuDeviceMemory.cu:170-209):
 * ssss900: the SS score table expanded to 30x30 (src/env/ssss);
 * nnw: SS18 network weights + adjustment parameters (93 floats,
 * CuDeviceMemory.cu:40-80 layout CuBatchProcessing_com.cuh:44-94) */
int  cbp_set_options(cbp_handle* h, const cbp_options* opt);
int  cbp_upload_tables(cbp_handle* h, const float* ssss900, const float* nnw, int32_t nnw_count);

/* the context-vector score map (src/env/cvs2s) as CVS2Scores reads it
 * (libpro/srcpro/CVS2Scores.cpp:54-466): the first table's keys and scores,
 * the scores smoothed by the centred five-point moving average of
 * SmoothScores, missing keys inserted at step 1/2, shift = number of
 * negative keys.  Host only, no GPU.  keys and values hold cap entries;
 * *card receives the table size */
int  cbp_cvs2s_read(const char* path, float* keys, float* values, int32_t cap,
                    int32_t* card, int32_t* shift);
/* upload the map for the CV score term (CuDeviceMemory::CacheCVS2ScoresDevice,
 * cupro/CuDeviceMemory.cu:671-723): a score x maps to values[0] when
 * x <= keys[0], to values[card-1] when x >= keys[card-1], else to
 * values[shift + rint(2x)] (SerializedCVS2Scores::GetScoreE1S1Step2Boundary,
 * cupro/SerializedCVS2Scores.cuh:377-410) */
int  cbp_upload_cvs2s(cbp_handle* h, const float* keys, const float* values, int32_t card,
                      int32_t shift);

/* data: queries (CuDeviceMemory::CacheData / CacheDataDevice,
 * cupro/CuDeviceMemory.cu:234-250, 1105-1141); DB chunk
 * (CuBatchProcessing::TransferCPMData -> CuDeviceMemory::TransferCPMData /
 * CopyCPMDataToDevice, cupro/CuDeviceMemory.cu:251, 1036-1104) */
int  cbp_upload_queries(cbp_handle* h, const cbp_profiles* q);
int  cbp_upload_chunk(cbp_handle* h, const cbp_profiles* db);
/* the reference's per-field buffers (char* arrays of pmv2DTotFlds begin and
 * end pointers, cupro/PM2DVectorFields.h:41-81), as CuBatchProcessing::
 * TransferCPMData(bdbCpmbeg, bdbCpmend, szCpm2dvfields) receives them for a
 * DB chunk (cupro/CuBatchProcessing.cuh:60-68): field f spans
 * [beg[f], end[f]) -- per profile: 20 background probabilities, ENO (float),
 * length (int), distance (unsigned, positions before the profile); per
 * profile position plus one beginning row: 7 log transitions; per position:
 * 20 target frequencies, 19 CV entries, CV prior, CV norm, 3 SS
 * probabilities, HDP probability, HDP index, 3 addresses (int), residue
 * (char), SS state (char).  CV, HDP and address fields may be empty; the
 * distances may start at any offset (they are re-based to the first
 * profile). */
#define CBP_PM_NFIELDS 81
/* profile and position counts of a field set (host only, no GPU) */
int  cbp_fields_count(const void* const* beg, const void* const* end, int32_t* nprof, int64_t* npos);
/* gather a field set into the field-major arrays of cbp_profiles, sized by
 * cbp_fields_count (host only, no GPU) */
int  cbp_fields_pack(const void* const* beg, const void* const* end,
                     float* bkg, float* eno, int32_t* len, uint32_t* dist,
                     float* trgfrq, float* ssprb, char* aa, char* sss, float* trans);
/* the context-vector fields of a field set, cv [19][npos], cvprior and
 * cvnorm2 [npos]; returns 1 when the set carries them (every CV field spans
 * npos values), 0 when the CV fields are empty (nothing written), <0 on
 * error (host only, no GPU) */
int  cbp_fields_pack_cv(const void* const* beg, const void* const* end,
                        float* cv, float* cvprior, float* cvnorm2);
/* cbp_upload_chunk of a field set (CuBatchProcessing::TransferCPMData,
 * cupro/CuBatchProcessing.cuh:60-68 -> CuDeviceMemory::CopyCPMDataToDevice,
 * cupro/CuDeviceMemory.cu:1036-1104) */
int  cbp_upload_chunk_fields(cbp_handle* h, const void* const* beg, const void* const* end);
/* a DB of two parts searched as one profile set: the cached part bdb1
 * (beg1, end1; may be NULL) followed by the chunk bdbC (beg2, end2), the
 * two parts ProcessScoreMatrix receives (cupro/CuBatchProcessing.cuh:82-94,
 * numbered as cudp/CuBatchDP_init_btck.cu:96-110 numbers them) */
int  cbp_upload_db_fields(cbp_handle* h, const void* const* beg1, const void* const* end1,
                          const void* const* beg2, const void* const* end2);
/* appends one query, given as the reference's querypmbeg field pointers
 * advanced to it (cualn/DevCommThread.cpp:312-345) and its length nqyposs,
 * to the query set; *qidx receives its index; the other queries keep their
 * indices and device preparations (the query argument of
 * CuBatchProcessing::ProcessScoreMatrix, cupro/CuBatchProcessing.cuh:82-94) */
int  cbp_add_query_fields(cbp_handle* h, const void* const* beg, uint64_t nqyposs, int32_t* qidx);

/* COMER text profiles ("COMER profile v2.0", as makepro writes them:
 * TextWriteProfileCondensed, libpro/srcpro/PMProfileModel.cpp:2295-2510)
 * read into a profile set the library owns (TextReadProfileHeader and
 * TextReadProfileData, cupro/IOProfileModel.cu:46-1021).  Host only, no
 * GPU.  A file holds one profile or several one after another; every
 * cbp_proset_read_text call appends its file's profiles.  The set's
 * cbp_profiles (valid until the next read or the free) feeds
 * cbp_upload_queries / cbp_upload_chunk; descriptions and file names are
 * kept per profile for the output. */
typedef struct cbp_proset cbp_proset;
int  cbp_proset_create(cbp_proset** out);
int  cbp_proset_read_text(cbp_proset* s, const char* path);
/* a binary profile database, format v2.2 ("COMER profile bin database
 * v2.2": Db::MakeBin_v2_2, libpro/srcpro/Db.cpp:333-383, read as
 * TdDataReader::OpenAndReadPreamble, cupro/TdDataReader.cpp:385-467, and the
 * chunk reads after it do), into an empty set: its sections are the
 * field-major arrays of cbp_profiles; context vectors always come with it */
int  cbp_proset_read_bin(cbp_proset* s, const char* path);
const cbp_profiles* cbp_proset_profiles(const cbp_proset* s);
const char* cbp_proset_desc(const cbp_proset* s, int32_t i);
const char* cbp_proset_name(const cbp_proset* s, int32_t i);
const char* cbp_proset_error(const cbp_proset* s);
void cbp_proset_free(cbp_proset* s);

/* whole-database size used for thresholds and E-values
 * (CuBatchProcessing::SetDbDetails) */
int  cbp_set_db_details(cbp_handle* h, uint64_t dblen, uint64_t nseqs);

/* score threshold and log E threshold for a query of length qlen searched
 * against dblen positions (JobDispatcher::CalculateScoreThreshold,
 * cualn/JobDispatcher.cpp:809-860) */
int  cbp_score_threshold(const cbp_options* opt, uint64_t qlen, uint64_t dblen,
                         float* scorethld, float* logevthld);

/* full per-(chunk, query) pipeline: score -> Viterbi -> statistics -> MAP
 * (CuBatchProcessing::BatchProcessScoreMatrixDevice, CuBatchProcessing.cu:229-572) */
int  cbp_process(cbp_handle* h, int32_t qidx, float scorethld, float logevthld,
                 cbp_results* out);

/* all listed queries against the uploaded chunk; thresholds per query from
 * cbp_score_threshold with the options' EVAL and the whole-database size
 * (the reference's DevCommThread loop over a query batch,
 * cualn/DevCommThread.cpp:296-356, calling ProcessScoreMatrix per query);
 * outs[i] stays valid until the next cbp_process* call on the handle */
int  cbp_process_batch(cbp_handle* h, const int32_t* qidx, int32_t nq,
                       cbp_results* outs /* [nq] */);

/* whole-database search of the listed queries in chunks (the reference's
 * chunk loop: JobDispatcher reads the DB chunk by chunk,
 * cualn/JobDispatcher.cpp:1103-1134, with TdDataReader's read-ahead,
 * cupro/TdDataReader.cpp:575-692; DevCommThread searches every query on each
 * chunk, cualn/DevCommThread.cpp:296-356; AlnWriter::MergeResults merges a
 * query's parts, cualn/AlnWriter.cpp:218-267).  db is the whole DB in host
 * memory (borrowed for the call); it is cut into chunks of whole profiles of
 * at most chunk_pos positions (0: half of cbp_chunk_limits' maximum for the
 * longest listed query, two chunks being resident); a copy thread uploads
 * chunk k+1 through pinned staging buffers while chunk k is searched; the
 * thresholds and statistics use the whole DB's size (cbp_set_db_details, or
 * db's own size when unset).  outs[i] is query qidx[i]'s result over the
 * whole DB, equal to a search of the DB as one chunk: rows in whole-DB
 * profile order, order by (log E, profile index), the first max_map_pairs
 * selected pairs drawn.  The handle's own chunk (cbp_upload_chunk) is
 * released for searching: upload a chunk again before cbp_process.
 * outs stay valid until the next cbp_search_db call on the handle. */
int  cbp_search_db(cbp_handle* h, const cbp_profiles* db, uint64_t chunk_pos,
                   const int32_t* qidx, int32_t nq, cbp_results* outs /* [nq] */);
/* number of chunks of the last cbp_search_db call */
int  cbp_last_search_chunks(const cbp_handle* h, int32_t* nchunks);

/* a query's results as the reference's plain-text output file (host only,
 * no GPU): the programme header, the query, the database lines, one
 * annotation per hit, the alignment records and the summary
 * (AlnWriter::WriteResultsPlain, cualn/AlnWriterPlain.cpp:19-121, with the
 * records of CuBatchProcessingFinalizer::CompressResultsPlain,
 * cupro/CuBatchProcessingFinalizerPlain.cpp:31-143).  Hits are the pairs
 * with log E <= res->logevthld and an alignment drawn, in res->order.  The
 * DB profile arrays are indexed by the records' profile index
 * (CBP_F_ORGPRONO).  out may be NULL to get the length; otherwise cap must
 * exceed it (the text is NUL-terminated). */
typedef struct {
    const char* progname;          /* programme name and version line (NULL: none) */
    const char* progversion;
    const char* const* references; /* NULL-terminated reference lines, or NULL */
    const char* query_desc;
    int32_t     query_len;
    float       query_eno;
    const char* query_sss;         /* [query_len] SS states (' ', 'e', 'h') or NULL */
    const char* const* db_names;   /* database names */
    int32_t     n_db_names;
    uint64_t    db_size;           /* total positions of the database */
    uint64_t    db_entries;        /* profiles of the database */
    const char* const* pro_desc;   /* per DB profile: description */
    const int32_t* pro_len;        /* per DB profile: length */
    const float* pro_eno;          /* per DB profile: ENO */
    const uint32_t* pro_dist;      /* per DB profile: first position (for pro_sss) */
    const char* pro_sss;           /* SS states of the DB positions, or NULL */
    int32_t     deltalen;          /* length adjustment (cbp_search_space) */
    int32_t     nohits, noalns;    /* NOHITS, NOALNS (options.txt:11,13; < 0: all) */
    int32_t     dsclen, dscwidth, alnwidth;   /* DSCLEN, DSCWIDTH, ALNWIDTH (:18,20,22) */
    int32_t     show;              /* SHOW (:15): the statistics footer of each record */
} cbp_plain_info;
int  cbp_format_plain(const cbp_options* opt, const cbp_plain_info* info, const cbp_results* res,
                      char* out, size_t cap, size_t* len);

/* the same results as the reference's JSON output file (host only):
 * {"comer_search": {program, version, references, query, database,
 * message, search_summary, search_hits, search_statistics}}
 * (AlnWriter::WriteResultsJSON, cualn/AlnWriterJSON.cpp:20-143, with the
 * records of CuBatchProcessingFinalizer::CompressResultsJSON,
 * cupro/CuBatchProcessingFinalizerJSON.cpp:35-382); descriptions escape
 * '"', '\\' and '/' (FormatDescriptionJSON, liblib/fmtdescription.h:96-126).
 * Takes the fields of cbp_plain_info that the JSON file uses (dscwidth,
 * alnwidth and show are not). */
int  cbp_format_json(const cbp_options* opt, const cbp_plain_info* info, const cbp_results* res,
                     char* out, size_t cap, size_t* len);

/* effective search space and length adjustment of a query of qlen
 * positions against dblen positions in nseqs profiles
 * (AbstractScoreMatrix::ComputeLengthAdjustment,
 * libpro/srcsco/AbstractScoreMatrix.cpp:247-390); host only */
int  cbp_search_space(const cbp_options* opt, uint64_t qlen, uint64_t dblen, uint64_t nseqs,
                      uint64_t* sspace, int32_t* deltalen);

/* ---- stage entry points (exposed for parity tests and profiling) ---- */

/* phase-1 fused score+Viterbi over all chunk profiles for query qidx;
 * keys[p] = (float bits of max MM score) << 32 | ~(x << 16 | y), x, y the
 * profile-local coordinates of the maximum (0 if no positive cell) */
int  cbp_phase1_keys(cbp_handle* h, int32_t qidx, uint64_t* keys /* [nprof] */);

/* FP32 score matrix S and its SS component for query qidx vs chunk profile
 * pronr, row-major [qlen][dblen] in host memory, mod may be NULL
 * (CalcSM_Init_SMEMUnroll2x + CalcSM_SSSS_SMEMUnroll2x,
 * cusco/CuBatchScoreMatrix.cu:118-377) */
int  cbp_score_block(cbp_handle* h, int32_t qidx, int32_t pronr, float* S, float* mod);

/* cbp_phase1_keys without any allocation (no reference counterpart; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid it).  Every device buffer the
 * call writes besides the uploaded chunk and query set -- the maxima, the
 * score weights, the stripe hand-off of multi-stripe queries and the
 * query's phase-1 tables -- lives in one static device array of 64 MiB that
 * each GPU the library runs on holds until the process ends; a query and
 * chunk that need more (88 B per profile, 20 B per position for queries of
 * more than one stripe of 64*R rows, about 350 B per query row) are refused
 * with CBP_ERR_ARG before any launch.  The array is shared by every call on
 * one GPU: calls that overlap on one GPU (two threads or two handles) race */
int  cbp_phase1_keys_noalloc(cbp_handle* h, int32_t qidx, uint64_t* keys /* [nprof] */);

/* tuning of the phase-1 launch geometry: target DB columns per wave segment
 * (>= 64), maximum query rows per lane (1..5) and waves per workgroup
 * (1, 2, 4, 8); no reference counterpart (the reference fixes 32x32 blocks,
 * cudp/CuBatchDP_com.h:16-30) */
int  cbp_set_tuning(cbp_handle* h, int32_t seg_cols, int32_t rmax, int32_t nw);

/* chunk limits for a query of qlen positions: the most DB positions and
 * profiles a chunk may hold, and the most positions and profiles of pass 2
 * (statistics and MAP), under the handle's memory budget (device free memory
 * when the budget is 0); replaces CuBatchProcessing::GetCurrentMaxDbPos,
 * GetCurrentMaxNDbPros and their Pass2 variants (cupro/CuBatchProcessing.cuh:106-111)
 * backed by CuDeviceMemory::CalcMaxDbDataChunkSize (cupro/CuDeviceMemory.cu:269-330) */
int  cbp_chunk_limits(cbp_handle* h, uint64_t qlen, uint64_t* maxpos, uint64_t* maxpros,
                      uint64_t* maxpos2, uint64_t* maxpros2);

/* autotuner: times the phase-1 launch of query qidx over segment widths
 * {2048, 4096, 8192, 16384} columns (tile width), {4, 8} waves per workgroup
 * and at most {5, 4, 3} query rows per lane (anti-diagonal depth), keeps the
 * fastest (cbp_set_tuning) and returns it (any out pointer may be NULL); no
 * reference counterpart */
int  cbp_autotune(cbp_handle* h, int32_t qidx, int32_t* seg_cols, int32_t* nw, int32_t* rmax);

/* capture the phase-1 stage of cbp_process / cbp_process_batch in a HIP graph
 * and replay it while the batch, chunk and tuning are unchanged (the first
 * run of a batch is eager); no reference counterpart */
int  cbp_set_graph(cbp_handle* h, int32_t on);

/* phase-1 sweeps and the candidates' backtrack pass of a cbp_process_batch
 * call of two queries or more from a device work queue (on = 1, the
 * default): persistent workgroups pull (query, segment group) tiles binned
 * by rows per lane, one launch per bin; on = 0 launches each query on its
 * own (query streams).  Results are identical either way; no reference
 * counterpart (the reference's DevCommThread runs one query at a time,
 * cualn/DevCommThread.cpp:296-356) */
int  cbp_set_queue(cbp_handle* h, int32_t on);

/* the phase-1 DB sweep (one-query launches and the batch work queue) forms
 * each cell's 20-term score sum on the matrix cores (on = 1) instead of the
 * vector ALUs (on = 0): every group of four lanes computes its rows against
 * four columns with v_mfma_f32_4x4x1_16b_f32 chained over the 20 terms, one
 * product added per instruction, which rounds as the fmaf chain does, so
 * keys and every later stage are bit-identical.  The backtrack pass and the
 * context-vector sweep keep the vector form.  Replaces the score-matrix
 * contraction of CalcSM_Init_SMEMUnroll2x (cusco/CuBatchScoreMatrix_Init.cu:45-186) */
int  cbp_set_mfma(cbp_handle* h, int32_t on);

/* MAP passes of a one-query call (the pairs no longer than half the longest
 * run as their own group beside the long pairs): rmax = 1..5 gives that
 * group at most rmax query rows per lane and as few stripes (waves per pair)
 * as that allows, so each wave works on more rows and fewer waves wait at
 * the band's per-step barriers; rmax = 0 keeps the long pairs' geometry
 * (one band of up to 8 waves, R = ceil(Lq/512)) for every pair.  Results are
 * bit-identical either way.  Replaces the fixed 32x32-block schedule of
 * ExecMAPDP_Parent_DynPlm (cumapdp/CuBatchMAPDP_prnt.cu:18-115) */
int  cbp_set_map_rows(cbp_handle* h, int32_t rmax);

/* MAP forward and backward passes (on = 1, the default): for queries whose
 * stripes fit one band of 8 waves, persistent workgroups stream their pairs
 * through the band back to back, so waves do not idle while a pair's band
 * fills and drains; on = 0 runs one workgroup per pair.  Results are
 * bit-identical either way.  Replaces the per-pair launch loops of
 * ExecMAPDP_Parent_DynPlm (cumapdp/CuBatchMAPDP_prnt.cu:18-115) */
int  cbp_set_map_pipe(cbp_handle* h, int32_t on);

/* stage lists built on the device (on = 1, the default): the candidate,
 * pass and drawing lists, the (log E, profile) order and the MAP plan of
 * cbp_process / cbp_process_batch are ordered compactions and sorts on the
 * device, so no host decision separates the stages and cbp_set_graph
 * captures the whole sequence (phase 1 to the records); a call in which a
 * list exceeds its fixed capacity runs again with host-built lists (on = 0
 * always does).  Results are identical either way.  The reference builds
 * its pass list on the device too (cudp/CuBatchDP_final.cu:250-264) and
 * schedules statistics and MAP from the device
 * (cuss/CuBatchSS_dynplm.cu:697-830, cumapdp/CuBatchMAPDP_prnt.cu:18-115) */
int  cbp_set_device_lists(cbp_handle* h, int32_t on);
/* capacities of the device lists per query: candidates, passed pairs (at
 * most 16384, the order is sorted in LDS), backtrack-code words and MAP
 * cells per stream slot (defaults 65536, 16384, 2^27, 2^28) */
int  cbp_set_device_list_caps(cbp_handle* h, int32_t cand, int32_t passed, int64_t code_words,
                              int64_t map_cells);
/* queries of the last cbp_process* call that ran again with host-built
 * lists because a device list exceeded its capacity */
int  cbp_last_fallbacks(const cbp_handle* h, int32_t* n);

/* device-side elapsed time of the last cbp_phase1_keys / cbp_process call's
 * dominant kernel (ms, HIP events on its stream) and the number of launches */
int  cbp_last_kernel_ms(const cbp_handle* h, double* ms, int32_t* nlaunch);

/* ---- differentiable soft DP (no reference counterpart) ----
 * The model is the MAP forward model of ExecMAPDP_Fwd_Unroll32x
 * (cumapdp/CuBatchMAPDP_fwd.cu:291-372) with exact log-sum-exp, over the
 * uploaded queries and chunk, plus five global log-offsets theta added to the
 * transition classes: insert open, insert extend, delete open, delete
 * extend, gap close.  logZ = log sum_ij exp(MM(i,j)).  The adjoint returns
 * dlogZ/dS, dlogZ/dtheta and, optionally, the MM posteriors
 * exp(F+B-logZ), the quantity of CuBatchMAPDP_bwd.cu:531-544.
 * S, theta, logZ, gZ, dS, dtheta, post and ws are device pointers; stream is
 * a hipStream_t (NULL: the handle's stream); calls are asynchronous. */
typedef struct {
    int32_t qidx;        /* query index in the uploaded query set */
    int32_t pronr;       /* profile index in the uploaded chunk */
    int64_t soff;        /* float offset of the pair's row-major [Lq][Ldb] block in S, dS, post */
} cbp_sdp_pair;
/* plan the pairs; ckpt = 0 keeps every forward cell for the adjoint, K > 0
 * keeps a checkpoint every K anti-diagonal steps and recomputes, -1 picks
 * K = ceil(sqrt(max steps)); K is rounded down to 16, 32 or 64 (at least 16);
 * returns the workspace size the caller allocates */
int  cbp_sdp_prepare(cbp_handle* h, const cbp_sdp_pair* pairs, int32_t npairs, int32_t ckpt,
                     int64_t* ws_bytes);
/* score matrices of the prepared pairs into S (device), the score of
 * CalcSM_Init_SMEMUnroll2x + CalcSM_SSSS_SMEMUnroll2x from the raw profile
 * fields (cusco/CuBatchScoreMatrix_Init.cu:45-186, CuBatchScoreMatrix_ssss.cu:51-280);
 * the context-vector term is not added here (a caller adds it to S) */
int  cbp_sdp_scores(cbp_handle* h, float* S, void* stream);
/* logZ[npairs] (natural log) */
int  cbp_sdp_forward(cbp_handle* h, const float* S, const float* theta, float* logZ,
                     void* ws, void* stream);
/* after cbp_sdp_forward with the same S, theta, ws: gZ [npairs] upstream
 * gradient (NULL = 1), dS [like S], dtheta [npairs][5], post (NULL or like S) */
int  cbp_sdp_adjoint(cbp_handle* h, const float* S, const float* theta, const float* logZ,
                     const float* gZ, float* dS, float* dtheta, float* post, void* ws, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* COMER2_HIP_H */
