/* ggcat_gpu.h -- C ABI of the MI355X (gfx950) implementation of GGCAT's
 * two data-parallel front stages: minimizer bucketing and per-sub-bucket
 * k-mer merging.
 *
 * Plain C types only; no function unwinds across this boundary.  Every call
 * returns a GGCAT_GPU_* status; on failure ggcat_gpu_last_error() describes
 * it.  One context owns one HIP stream and its device buffers; calls on one
 * context must not overlap (the reference's API is not re-entrant either,
 * crates/api/src/lib.rs:118-141).
 *
 * Input layout ("joined"): all records concatenated, each followed by at
 * least one byte that is not A/C/G/T (e.g. '\n').  Non-ACGT bytes end an
 * ACGT run exactly like 'N' after the reference's normalisation
 * (crates/io/src/sequences_reader.rs:26-54,
 *  crates/minimizer_bucketing/src/sequences_splitter.rs:15-40).
 */
#ifndef GGCAT_GPU_H
#define GGCAT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ggcat_gpu_ctx ggcat_gpu_ctx;

enum {
    GGCAT_GPU_OK = 0,
    GGCAT_GPU_EINVAL = 1, /* bad argument */
    GGCAT_GPU_EHIP = 2,   /* HIP runtime error */
    GGCAT_GPU_ENOMEM = 3, /* device allocation failed */
    GGCAT_GPU_ESTATE = 4, /* call out of order (e.g. merge before bucketing) */
    GGCAT_GPU_EAGAIN = 5  /* nothing ready yet (serving poll) */
};

enum { GGCAT_GPU_HASH_SEQ = 0, GGCAT_GPU_HASH_RK128 = 1 };

/* Parameters of one build.  Mirrors the arguments run_assembler derives
 * (crates/assembler/src/lib.rs:82-155): k, m (--minimizer-length), the
 * bucket count log b, the second-bucket count log sb
 * (crates/io/src/lib.rs:113-128), forward_only (-f), the merging hash
 * (-w, crates/api/src/utils.rs:10-68) and min_multiplicity (-s). */
typedef struct {
    uint32_t k;
    uint32_t m;
    uint32_t buckets_log;
    uint32_t second_buckets_log;
    uint32_t forward_only;
    uint32_t hash_kind; /* GGCAT_GPU_HASH_SEQ or GGCAT_GPU_HASH_RK128 */
    uint32_t min_multiplicity;
} ggcat_gpu_params;

/* Creates a context on HIP device `device`.
 * Replaces GGCATInstance::create (crates/api/src/lib.rs:141-215) for the
 * device side of the two stages; returns GGCAT_GPU_OK and *out, or an error
 * with *out == NULL (the reference's ggcat_create returns null,
 * crates/capi/src/lib.rs:17-55). */
int ggcat_gpu_create(int device, ggcat_gpu_ctx** out);

/* Frees the context and every device buffer it owns. */
void ggcat_gpu_destroy(ggcat_gpu_ctx* ctx);

/* Description of the last failure on this context (never NULL). */
const char* ggcat_gpu_last_error(const ggcat_gpu_ctx* ctx);

/* Minimizer bucketing of one joined input.
 * Replaces minimizer_bucketing
 * (crates/assembler_minimizer_bucketing/src/lib.rs:279-340), called by
 * run_assembler at crates/assembler/src/lib.rs:157-173.  Produces, in device
 * memory owned by the context, every super-k-mer record grouped by
 * sub-bucket (bucket << sb | second_bucket; bucket 2^b is the duplicates
 * bucket, crates/assembler/src/lib.rs:147-155).
 *   data, n_bytes   joined input; a host pointer, or a device pointer when
 *                   input_on_device != 0 (then it must stay valid until the
 *                   call returns and be 16-byte aligned)
 *   color_starts, colors, n_colors
 *                   optional per-record colors (colored builds): record
 *                   color ranges start at the sorted byte offsets
 *                   color_starts[i] (color_starts[0] == 0); NULL/0 when
 *                   uncolored.  Host pointers. */
int ggcat_gpu_bucketing(ggcat_gpu_ctx* ctx, const uint8_t* data, uint64_t n_bytes, int input_on_device,
                        const uint64_t* color_starts, const uint32_t* colors, uint32_t n_colors,
                        const ggcat_gpu_params* params);

/* Sub-buckets of the last bucketing that the reference would resplit as
 * outliers (crates/kmers_transform/src/lib.rs:171-371): buckets in the
 * reference's processing order (largest first after the min(normal / 8,
 * 2 * threads) smallest, then a large bucket alternated with small ones
 * matching its size), a running average of records per non-empty sub-bucket
 * (at least 5000), and a sub-bucket above 8x it resplit into next_pow2(records
 * / average) buckets clamped to [4, 1024] (crates/config/src/lib.rs:62-66).
 * The order uses the serialized record bytes of each bucket and the records
 * as bucketed (the reference uses compressed file sizes and compacted
 * counts).  Writes up to cap entries (sub-bucket, resplit bucket count,
 * records); *n_outliers = all of them.  Detection only: without
 * ggcat_gpu_resplit_outliers the merge treats outliers like every other
 * sub-bucket. */
int ggcat_gpu_outliers(ggcat_gpu_ctx* ctx, uint32_t threads, uint32_t* n_outliers, uint32_t* sub, uint32_t* resplit,
                       uint64_t* records, uint32_t cap);

/* Resplits the outliers of the last bucketing (those ggcat_gpu_outliers
 * finds with `threads`) as crates/kmers_transform/src/resplitter.rs:47-183
 * does before their merge: every stored record of an outlier is split again
 * by process_sequence (crates/assembler_minimizer_bucketing/src/lib.rs:
 * 152-270) with m' = k - 10 (min(m, 2) for k <= 11,
 * crates/assembler_kmers_merge/src/lib.rs:236-246), into the outlier's
 * resplit bucket count of normal buckets plus a duplicates bucket, keeping
 * its multiplicity and color, with its include flags as the new ends'.
 * Outlier i's buckets become sub-buckets first_sub[i] .. first_sub[i] +
 * resplit[i] (the duplicates bucket last), numbered from the bucketing's
 * sub-bucket count in outlier order; the next ggcat_gpu_kmers_merge counts
 * each as a sub-bucket of its own (the reference runs the same MapEntry
 * rules on them).  Writes up to cap entries; *n_outliers = all of them.
 * Once per bucketing; ggcat_gpu_records_serialize and ggcat_gpu_outliers
 * then return GGCAT_GPU_ESTATE. */
int ggcat_gpu_resplit_outliers(ggcat_gpu_ctx* ctx, uint32_t threads, uint32_t* n_outliers, uint32_t* sub,
                               uint32_t* resplit, uint32_t* first_sub, uint32_t cap);

/* The resplit of ggcat_gpu_resplit_outliers applied to the n listed
 * sub-buckets of the last bucketing (distinct, below the bucketing's
 * sub-bucket count) with count[i] normal buckets each (a power of two in
 * [1, 1024]) instead of the detected outliers: sub-bucket i's buckets become
 * first_sub[i] .. first_sub[i] + count[i].  Once per bucketing. */
int ggcat_gpu_resplit(ggcat_gpu_ctx* ctx, uint32_t n, const uint32_t* sub, const uint32_t* count,
                      uint32_t* first_sub);

/* Sizes of the last bucketing result: records, sub-buckets (resplit
 * buckets included), and the bytes needed by ggcat_gpu_records_copy for the
 * bases buffer. */
int ggcat_gpu_records_info(ggcat_gpu_ctx* ctx, uint64_t* n_records, uint32_t* n_sub, uint64_t* n_bases);

/* Decoded copy of the last bucketing result (test/inspection path):
 * per record its sub-bucket, length, flags (READ_FLAG_INCL_BEGIN=1,
 * READ_FLAG_INCL_END=2 as stored), minimizer position, color, offset into
 * `bases`, and the bases as ASCII (already reverse-complemented when the
 * reference stores them reverse-complemented,
 * crates/assembler_minimizer_bucketing/src/lib.rs:228-263).
 * Any output pointer may be NULL. */
int ggcat_gpu_records_copy(ggcat_gpu_ctx* ctx, uint32_t* sub, uint32_t* len, uint8_t* flags, uint16_t* minpos,
                           uint32_t* color, uint64_t* off, uint8_t* bases);

/* Serializes the last bucketing result in the byte layout of the
 * reference's bucket files, so that its unchanged downstream can read it:
 * CompressedReadsBucketDataSerializer::write_to
 * (crates/io/src/concurrent/temp_reads/creads_utils.rs:374-432) as
 * instantiated by the minimizer bucketing
 * (crates/minimizer_bucketing/src/lib.rs:271-283: WithSecondBucket,
 * NoMultiplicity, AssemblerMinimizerPosition, two flag bits, min_size = k).
 * Per record: the second-bucket byte; varint_flags(((len - k) <<
 * log2(next_pow2(k))) | minimizer position, flags)
 * (crates/io/src/compressed_read.rs:435-467, crates/io/src/varint.rs:26-57);
 * the bases packed 4 per byte, base i in bits 2(i mod 4) of byte i / 4, as
 * stored (crates/io/src/compressed_read.rs:610-633); for colored builds the
 * color as a varint, 0 when it repeats the previous record's, else color + 1
 * (crates/colors/src/parsers/separate.rs:98-115).  Bucket b's stream (its
 * 2^sb sub-buckets; the duplicates bucket is the last) is
 * out[bucket_off[b] .. bucket_off[b + 1]); bucket_off holds
 * n_sub / 2^sb + 1 entries.  The record order inside a stream is the
 * device's (the reference's depends on thread timing).  out == NULL: only
 * *n_bytes and bucket_off are filled.  Host buffers; out needs cap >=
 * *n_bytes. */
int ggcat_gpu_records_serialize(ggcat_gpu_ctx* ctx, uint8_t* out, uint64_t cap, uint64_t* bucket_off,
                                uint64_t* n_bytes);

/* Per-sub-bucket canonical k-mer counting of the last bucketing result.
 * Replaces kmers_merge (crates/assembler_kmers_merge/src/lib.rs:159-284),
 * called at crates/assembler/src/lib.rs:269-319, for its k-mer table: every
 * k-mer whose multiplicity (crates/structs/src/map_entry.rs:79-83) is
 * >= params.min_multiplicity is kept, with the MapEntry flags. */
int ggcat_gpu_kmers_merge(ggcat_gpu_ctx* ctx);

/* Number of solid (sub-bucket, k-mer) entries of the last merge. */
int ggcat_gpu_solid_info(ggcat_gpu_ctx* ctx, uint64_t* n_solid, uint32_t* key_words);

/* Copy of the solid entries: keys (key_words uint64 per entry, low word
 * first), multiplicity, flags and sub-bucket.  Order is unspecified. */
int ggcat_gpu_solid_copy(ggcat_gpu_ctx* ctx, uint64_t* keys, uint32_t* multiplicity, uint8_t* flags,
                         uint32_t* sub);

/* Color sets of the solid entries of a colored build, in ggcat_gpu_solid_copy
 * order: color_words uint32 per entry, bit c set when color c occurs among
 * the k-mer's occurrences -- the sorted unique union of
 * crates/colors/src/managers/multiple.rs:99-208 (add_temp_buffer_structure_el,
 * process_colors) before its subset id is assigned.  Color ids are < 1024.
 * bitmaps may be NULL to query *color_words. */
int ggcat_gpu_solid_colors(ggcat_gpu_ctx* ctx, uint32_t* color_words, uint32_t* bitmaps);

/* Schedule knobs of the merge (no counterpart in the reference, whose
 * schedule is its thread pool).  They change which kernel and table size
 * handle a work item, never the result.  Zero selects the default.
 *   item_occ          target k-mer occurrences per work item (runs of
 *                     physical units of one sub-bucket); default: half the
 *                     LDS table (1024 for 64-bit keys, 512 for 128-bit)
 *   lds_max_occ       items with more occurrences skip the whole-item pass
 *                     and go straight to hash-partitioned passes; default
 *                     16x the LDS table
 *   merge_wgs_per_cu  persistent merge workgroups per CU; default 6
 *   part_occ          target occurrences per partition of an item that is
 *                     split into hash-partitioned passes; default half the
 *                     LDS table (a partition that still overflows is split
 *                     4 ways again)
 *   lds_table_log     log2 of the LDS table entries per workgroup (8..13;
 *                     default 11 for 64-bit keys, 10 for 128-bit; reduced
 *                     to fit 48 KiB with colors)
 *   table_pct         table slots per 100 occurrences of an item, rounded up
 *                     to a power of two (default 200: never overflows when
 *                     every k-mer is distinct; lower values trade overflow
 *                     redos for smaller tables)
 *   no_dedup          1: insert every record's k-mers even when identical
 *                     records could be counted once; 2: count identical
 *                     records once but let one thread roll each distinct
 *                     record's k-mers (the mode of 128-bit keys) instead of
 *                     spreading single k-mer positions over the threads;
 *                     3: hash the records of every work item, also of those
 *                     whose units the bucketing already deduplicated (no
 *                     unit above the deduplication's 1024-record limit),
 *                     which otherwise take their records by position
 *                     (no mode changes a result)
 *   unit_bytes        bucketing: input bytes per physical unit (sub-bucket
 *                     refinement); default 2000, 1000 for colored builds
 *   use_graphs        1: capture the bucketing staging pass (clears, stage
 *                     and deferred-record kernels) in a HIP graph per
 *                     argument set and replay it on later calls
 *   unit_cap          bucketing: slots of each unit's region, where records
 *                     are written as they are counted (default 2.5x the
 *                     records expected per unit); records past it go to a
 *                     spill list; the result never depends on it
 *   rec_cap           merge: identical-record table slots (power of two in
 *                     [16, 512], default 512); records are deduplicated in
 *                     chunks of rec_cap / 2 (at most one record per thread)
 *   merge_threads     threads per merge workgroup: 128 or 256 (default)
 *   read_block        input files: the read block of the reference's line
 *                     reader (0 = its DEFAULT_OUTPUT_BUFFER_SIZE, 4 MiB, else
 *                     at least 256); results are those of the reference run
 *                     with that block size (it decides where long FASTA
 *                     records are cut), so only tests set it
 *   host_reader       1: read input files on the host even when they could be
 *                     read into HBM and parsed there (same records)
 *   upload_chunk      host input of ggcat_gpu_bucketing: bytes per upload
 *                     chunk of the pipeline that stages resident tiles while
 *                     later chunks upload (0 = 64 MiB, the pinned staging
 *                     buffer; at most that); input shorter than two chunks
 *                     is copied whole before the staging pass */
typedef struct {
    uint32_t item_occ;
    uint32_t lds_max_occ;
    uint32_t merge_wgs_per_cu;
    uint32_t part_occ;
    uint32_t lds_table_log;
    uint32_t table_pct;
    uint32_t no_dedup;
    uint32_t unit_bytes;
    uint32_t use_graphs;
    uint32_t unit_cap;
    uint32_t rec_cap;
    uint32_t merge_threads;
    uint32_t read_block;
    uint32_t host_reader;
    uint32_t upload_chunk;
} ggcat_gpu_tuning;
int ggcat_gpu_set_tuning(ggcat_gpu_ctx* ctx, const ggcat_gpu_tuning* tuning);

/* Device-side timings of the last calls, in milliseconds:
 * [0] bucketing staging pass (with the uploads of chunked host input), [1] scan, [2] bucketing deduplication + region compaction,
 * [3] merge (whole-item pass), [4] merge (partitioned passes),
 * [5] whole bucketing call, [6] whole merge call, [7] unitig kernels,
 * [8] reading the input of the last build_from_files (wall clock: files
 * read into HBM and parsed there, or read on the host). */
int ggcat_gpu_timings(ggcat_gpu_ctx* ctx, double* ms, uint32_t n);

/* Counters of the last calls: [0] input bytes, [1] records,
 * [2] sum of record lengths (bases), [3] k-mer occurrences,
 * [4] work items split into hash-partitioned passes, [5] solid entries,
 * [6] k-mers shared between two units of a sub-bucket (combined on the
 * host after the merge), [7] merge kernel passes (1 + partition rounds),
 * [8] how the last build_from_files read its input (1: into HBM, parsed on
 * the device; 2: host reader), [9] long FASTA records cut on the device,
 * [10] records stored after identical copies in a unit were counted once
 * (the arena's slots; [1] counts every copy; copies past their unit's
 * region are stored apart, so this count can vary by a few records between
 * runs while every result stays the same). */
int ggcat_gpu_counters(ggcat_gpu_ctx* ctx, uint64_t* out, uint32_t n);

/* Reads FASTA/FASTQ files (optionally gzip-compressed) into the joined layout.
 * Restates crates/io/src/sequences_reader.rs:56-241 including long-record
 * chunking with a (k-1)-base copy-back.  With colors != 0, each file gets the
 * color equal to its index and *color_starts / *colors describe the ranges.
 * Buffers are allocated by the library; release them with
 * ggcat_gpu_free_host.  *estimated_bases is the reference's input size
 * estimate (crates/io/src/sequences_stream/fasta.rs:12-30).  Host-only: ctx
 * may be NULL (then ggcat_gpu_last_error(NULL) describes a failure). */
int ggcat_gpu_read_files(ggcat_gpu_ctx* ctx, const char* const* paths, uint32_t n_paths, uint32_t k, int colors,
                         uint8_t** joined, uint64_t* n_bytes, uint64_t** color_starts, uint32_t** color_ids,
                         uint32_t* n_colors, uint64_t* estimated_bases);
void ggcat_gpu_free_host(void* p);

/* Build options, mirroring ggcat_build_from_files
 * (crates/capi/src/lib.rs:567-599): m == 0 selects compute_best_m(k)
 * (crates/utils/src/lib.rs:29-40); buckets_log == 0 selects the automatic
 * bucket count (crates/io/src/lib.rs:67-140); the second-bucket count is
 * always automatic, as in run_assembler. */
typedef struct {
    uint32_t k;
    uint32_t m;
    uint32_t min_multiplicity;
    uint32_t buckets_log;
    uint32_t forward_only;
    uint32_t hash_kind;
    uint32_t colors;
} ggcat_gpu_build_options;

/* Parameters a build uses for inputs of `estimated_bases` bases (the sum of
 * the files' estimates): run_assembler's derivation,
 * crates/assembler/src/lib.rs:110-113 with crates/io/src/lib.rs:67-140, and
 * the m default of crates/api/src/lib.rs:291.  Host-only; needs no GPU. */
int ggcat_gpu_plan_build(const ggcat_gpu_build_options* options, uint64_t estimated_bases, ggcat_gpu_params* out);

/* Reads the files and runs bucketing and the k-mer merge: the work of
 * GGCATInstance::build_graph (crates/api/src/lib.rs:217-322) up to the end of
 * kmers_merge.  Results stay in the context (ggcat_gpu_records_* /
 * ggcat_gpu_solid_*); *used (optional) receives the derived parameters. */
int ggcat_gpu_build_from_files(ggcat_gpu_ctx* ctx, const char* const* paths, uint32_t n_paths,
                               const ggcat_gpu_build_options* options, ggcat_gpu_params* used);

/* ---- unitig chains inside each sub-bucket -------------------------------
 * Replaces compute_unitigs / try_extend_function
 * (crates/assembler_kmers_merge/src/unitigs_extender/hashmap.rs:162-297,
 * 442-601) on the last merge's solid table (seq-hash keys).  Every unitig is
 * a maximal path of k-mers of one sub-bucket that are each other's only
 * neighbour, not continuing through a k-mer whose flags are exactly
 * INCL_BEGIN or INCL_END (a continuation into another bucket); output as
 * ASCII, each unitig followed by '\n', with its sub-bucket and flags
 * (bit 0: the start continues in another bucket, bit 1: the end does,
 * bit 2: circular).  Unitigs are written from the end with the smaller
 * (key, side); cycles from their smallest k-mer, forward. */
int ggcat_gpu_unitigs(ggcat_gpu_ctx* ctx, uint64_t* n_unitigs, uint64_t* n_bytes);
int ggcat_gpu_unitigs_copy(ggcat_gpu_ctx* ctx, uint8_t* seq, uint32_t* sub, uint8_t* flags);
/* In/out extension flags of every solid entry (ggcat_gpu_solid_copy order):
 * bits 0-1 = number of solid k-mers of the same sub-bucket that can precede
 * it in its canonical orientation (2 = two or more), bits 2-3 = that can
 * follow it.  Valid after ggcat_gpu_unitigs. */
int ggcat_gpu_extension_copy(ggcat_gpu_ctx* ctx, uint8_t* ext);
/* Where the reference's merge sends each per-sub-bucket unitig of the last
 * ggcat_gpu_unitigs call (ParallelKmersMergeFinalExecutor::output_sequence,
 * crates/assembler_kmers_merge/src/final_executor.rs:103-246), in
 * ggcat_gpu_unitigs_copy order; any output may be NULL:
 *   kind        0 lonely unitig (no open end: straight to the output),
 *               1 lonely circular unitig, 2 partial unitig (an open end)
 *   bucket      partial unitigs: the result bucket, the smaller of the
 *               seq-hash buckets (get_bucket(0, buckets_log, ...)) of its
 *               open ends' k-mers; 0xFFFF otherwise
 *   rc          1 when the unitig is stored reverse-complemented
 *   last_align  the stored minimizer position: (length - k) % 4 when the
 *               hashed k-mer ends the stored sequence, else 0
 *   flags       HASH_ENDING (2) | OTHER_END (4)
 *               (crates/io/src/partial_unitigs_extra_data.rs:16-18)
 * buckets_log is the build's bucket count log (the result buckets have the
 * bucketing's count). */
int ggcat_gpu_partial_unitigs(ggcat_gpu_ctx* ctx, uint32_t buckets_log, uint8_t* kind, uint16_t* bucket, uint8_t* rc,
                              uint8_t* last_align, uint8_t* flags);

/* ---- maximal unitigs (partial unitigs joined across buckets) -------------
 * Replaces build_final_unitigs (crates/assembler_pipeline/src/lib.rs:111,
 * whose extend_unitigs at crates/assembler_pipeline/src/extend_unitigs.rs:
 * 348-993 joins the partial unitigs of all buckets; called from
 * crates/assembler/src/lib.rs:339-349) for uncolored seq-hash builds: the
 * maximal unitigs of the whole solid k-mer set of the last merge.  A k-mer
 * held by two sub-buckets (a boundary k-mer) is one node; every other rule is
 * the one of ggcat_gpu_unitigs with no side flagged.  Results replace the
 * previous unitig result and are read with ggcat_gpu_unitigs_copy (sub-bucket
 * 0 and flags 0 for every unitig, 4 for a circular one) and
 * ggcat_gpu_extension_copy (degrees over the whole set, for the first solid
 * entry of each k-mer; 0 for its other entries). */
int ggcat_gpu_maximal_unitigs(ggcat_gpu_ctx* ctx, uint64_t* n_unitigs, uint64_t* n_bytes);
/* Writes the last unitig result as FASTA, one record per unitig,
 * ">{id} LN:i:{length}" then the sequence on one line
 * (crates/sequence_output/src/structured_sequences/fasta.rs:153-200; ids
 * count from 0 in result order).  For colored builds the header continues
 * with the unitig's k-mers' color subset ids as runs " C:{id hex}:{count}"
 * in sequence order (crates/colors/src/managers/multiple.rs:707-740), ids as
 * ggcat_gpu_colormap_write writes them.  *n_written (optional) = records
 * written. */
int ggcat_gpu_unitigs_write_fasta(ggcat_gpu_ctx* ctx, const char* path, uint64_t* n_written);

/* ---- color map file (<graph>.colors.dat) ---------------------------------
 * The reference's run-length color map (magic "GGCAT_CMAP_RNLEN", version 1):
 * header, LZ4 frame of the color names, the color subsets in id order in LZ4
 * frames of 20000, and an index of the frames
 * (crates/colors/src/storage/serializer.rs:17-241, run_length.rs:26-334,
 * deserializer.rs:1-181).  Subset ids of a colored build are assigned to
 * the distinct color sets of its solid table in bitmap order (the reference
 * assigns them first-come, crates/colors/src/colors_memmap_writer.rs:50-59,
 * so its ids differ between runs while the sets do not). */

/* Subset id of every solid entry, in ggcat_gpu_solid_copy order (NULL to get
 * only *n_subsets).  Colored builds only. */
int ggcat_gpu_color_subsets(ggcat_gpu_ctx* ctx, uint64_t* n_subsets, uint32_t* subset_of_entry);
/* Writes the color map of the last colored build: color_names[i] names color
 * id i (n_names >= the build's largest color id + 1); *n_subsets (optional)
 * = subsets written.  Replaces the ColorsMemMapWriter of run_assembler
 * (crates/assembler/src/lib.rs:130-136). */
int ggcat_gpu_colormap_write(ggcat_gpu_ctx* ctx, const char* path, const char* const* color_names, uint32_t n_names,
                             uint64_t* n_subsets);
/* Writes a color map of explicit subsets (host only, no context): subset s
 * holds colors[subset_off[s] .. subset_off[s + 1]), strictly increasing and
 * below n_names.  Errors: ggcat_gpu_last_error(NULL). */
int ggcat_gpu_colormap_write_sets(const char* path, const char* const* color_names, uint32_t n_names,
                                  const uint64_t* subset_off, const uint32_t* colors, uint64_t n_subsets);
/* Reads a whole color map (host only; errors: ggcat_gpu_last_error(NULL)).
 * Mirrors ColorsDeserializer::new + get_color_mappings. */
typedef struct ggcat_gpu_colormap ggcat_gpu_colormap;
int ggcat_gpu_colormap_open(const char* path, ggcat_gpu_colormap** out);
int ggcat_gpu_colormap_info(const ggcat_gpu_colormap* map, uint32_t* n_colors, uint64_t* n_subsets);
/* NULL when color is out of range; valid until ggcat_gpu_colormap_close. */
const char* ggcat_gpu_colormap_color_name(const ggcat_gpu_colormap* map, uint32_t color);
/* Colors of one subset, ascending: *n always set, colors (cap entries) may be NULL. */
int ggcat_gpu_colormap_subset(const ggcat_gpu_colormap* map, uint64_t subset, uint32_t* colors, uint32_t cap,
                              uint32_t* n);
void ggcat_gpu_colormap_close(ggcat_gpu_colormap* map);

/* ---- queries ------------------------------------------------------------
 * Replaces the counting part of run_query (crates/querier/src/lib.rs:54-200):
 * the graph is the solid k-mer table of the last merge (the k-mer set of the
 * unitigs the reference would query), kept resident in HBM with its color
 * sets.  Per query (records of a joined buffer): query_kmers = length - (k-1)
 * saturating (crates/querier/src/lib.rs:138-150); matched = occurrences of
 * the query's all-ACGT k-mers present in the graph
 * (crates/querier/src/pipeline/parallel_kmers_query.rs:437-471); colored:
 * color_counts[q * n_colors + c] = matched k-mers whose color set holds c
 * (crates/querier/src/pipeline/colored_query_output.rs:200-320).  The CSV /
 * JSONL formatting of those numbers is host code (ggcat_amd/query.py). */

/* Builds the query index over the last merge's solid table (count, scan,
 * scatter into hash buckets).  Invalidated by the next bucketing call. */
int ggcat_gpu_index_build(ggcat_gpu_ctx* ctx);

/* Colors per query of the current index (0 when uncolored): one more than the
 * largest color id given to the colored build. */
int ggcat_gpu_index_colors(ggcat_gpu_ctx* ctx, uint32_t* n_colors);

/* Synchronous batch query.  queries: joined layout with exactly n_queries
 * records (host memory).  query_kmers and matched hold n_queries entries;
 * color_counts (colored index, may be NULL) n_queries * n_colors. */
int ggcat_gpu_query(ggcat_gpu_ctx* ctx, const uint8_t* queries, uint64_t n_bytes, uint32_t n_queries,
                    uint64_t* query_kmers, uint64_t* matched, uint32_t* color_counts);

/* Queries a graph given as files.  Mirrors ggcat_query_graph
 * (crates/capi/src/lib.rs:635-651, crates/api/src/lib.rs:325-382,
 * crates/querier/src/lib.rs:54-199): input_graph is a unitig FASTA (plain or
 * compressed); with colors != 0 its headers carry " C:{subset hex}:{count}"
 * runs and the color map is input_graph with its extension replaced by
 * "colors.dat".  Every k-mer of the unitigs is a graph k-mer (hash_kind as in
 * ggcat_gpu_params, canonical).  input_query is a FASTA/FASTQ file.  Writes
 * output_file_prefix (".csv" / ".jsonl" added when it has no extension;
 * compressed by its extension: lz4, gz, zst, bz2, xz) as the reference does:
 * CSV query_index,matched_kmers,query_kmers,match_percentage, or one JSON line
 * per query with matches, colors as numbers (color_output_format 0) or names
 * (1).  The output path is copied to out_path (cap bytes, may be NULL).
 * threads_count and minimizer_length are accepted and not needed;
 * forward_only must be 0.  The graph replaces the context's table (a later
 * ggcat_gpu_query queries it). */
int ggcat_gpu_query_graph(ggcat_gpu_ctx* ctx, const char* input_graph, const char* input_query,
                          const char* output_file_prefix, uint32_t k, uint32_t threads_count, int forward_only,
                          uint32_t minimizer_length, int colors, uint32_t color_output_format, uint32_t hash_kind,
                          char* out_path, uint64_t out_path_cap);

/* ---- serving: continuous batching of query requests ---------------------
 * No counterpart in the reference (each ggcat query call re-buckets the whole
 * graph).  Requests are admitted while their device and pinned staging
 * memory fits in `memory_budget` bytes (allocated once by serve_open; no
 * allocation happens per request), each on one of `streams` HIP streams, so
 * admitted requests overlap on the device; a request that does not fit waits
 * (first come, first admitted) until earlier ones are fetched. */
int ggcat_gpu_serve_open(ggcat_gpu_ctx* ctx, uint32_t streams, uint64_t memory_budget);
/* Queues one request (copies the query bytes); returns its ticket at once. */
int ggcat_gpu_serve_submit(ggcat_gpu_ctx* ctx, const uint8_t* queries, uint64_t n_bytes, uint32_t n_queries,
                           uint64_t* ticket);
/* Admits waiting requests that now fit and reports one finished ticket;
 * GGCAT_GPU_EAGAIN when none is finished (with block != 0 it waits for the
 * oldest admitted request instead). */
int ggcat_gpu_serve_poll(ggcat_gpu_ctx* ctx, int block, uint64_t* ticket);
/* Copies a finished request's results (as ggcat_gpu_query) and releases its
 * memory. */
int ggcat_gpu_serve_fetch(ggcat_gpu_ctx* ctx, uint64_t ticket, uint64_t* query_kmers, uint64_t* matched,
                          uint32_t* color_counts);
/* Requests submitted and not yet fetched, and those admitted to the device. */
int ggcat_gpu_serve_stats(ggcat_gpu_ctx* ctx, uint32_t* outstanding, uint32_t* admitted);
int ggcat_gpu_serve_close(ggcat_gpu_ctx* ctx);

/* Device-memory plumbing for callers that keep inputs resident in HBM (no
 * counterpart in the reference, which has no device memory).  The pointer
 * returned by ggcat_gpu_device_alloc is 256-byte aligned and can be passed to
 * ggcat_gpu_bucketing with input_on_device = 1. */
int ggcat_gpu_device_alloc(ggcat_gpu_ctx* ctx, uint64_t bytes, void** out);
int ggcat_gpu_device_free(ggcat_gpu_ctx* ctx, void* p);
int ggcat_gpu_copy_to_device(ggcat_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Device-to-device copy on the context's stream with a plain copy kernel
 * (16-byte aligned pointers and size); returns when the copy is done. */
int ggcat_gpu_copy_device(ggcat_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Synchronous device-to-host copy. */
int ggcat_gpu_copy_to_host(ggcat_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* ---- query batch that allocates nothing ----------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (tight serving loops, code that must not call hipMalloc).  Same
 * results as ggcat_gpu_query, but the queries (joined layout, n_queries
 * records each ending with '\n') and the outputs are device pointers owned
 * by the caller: d_query_kmers (n_queries entries, may be NULL), d_matched
 * (n_queries) and, for a colored index, d_color_counts (n_queries *
 * n_colors).  Everything else it needs (work cursor, newline counts, query
 * offsets) lives in a fixed 4 MiB device array of the library, one per GPU,
 * whose address is looked up by ggcat_gpu_create; the call makes no
 * allocation.  Limits: at most 500000 queries and 256 MiB of query bytes per
 * call (GGCAT_GPU_EINVAL beyond).  Calls that overlap on one GPU share that
 * array and race: serialise them.  Returns when the results are written. */
int ggcat_gpu_query_device_noalloc(ggcat_gpu_ctx* ctx, const uint8_t* d_queries, uint64_t n_bytes,
                                   uint32_t n_queries, uint64_t* d_query_kmers, uint64_t* d_matched,
                                   uint32_t* d_color_counts);

#ifdef __cplusplus
}
#endif

#endif /* GGCAT_GPU_H */
