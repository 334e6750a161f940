// Launchers of the HIP kernels (host-callable, stream-ordered, no syncs).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace sw {

// COUNT: total pairs; ANY: existence + smallest witness; FLAGS: bitmap over
// tile sets (contains some stream set); SCOUNT: pairs per stream set; SFLAGS:
// bitmap over stream sets (inside some tile set).
enum VsubMode { VSUB_COUNT = 0, VSUB_ANY = 1, VSUB_FLAGS = 2, VSUB_SCOUNT = 3, VSUB_SFLAGS = 4 };

// Parameters of the list-vs-list subset kernel (k_subset.hip).
struct VsubParams {
  const uint64_t* stream = nullptr;  // S_i, row-major W words
  uint64_t s_count = 0;
  const uint64_t* tile = nullptr;    // T_j, row-major W words
  uint64_t t_count = 0;
  uint32_t n = 0;                    // number of states
  int complement_tile = 0;           // test S_i against the complement of T_j
  int complement_stream = 0;         // use the complement of S_i
  int mode = VSUB_COUNT;
  int early_exit = 1;                // bit-walking kernel: stop a stream set once no lane can hit
  int ix_early_exit = 0;             // indexed kernel: the same, checked after every four states
  int proper_by_size = 0;            // skip pairs of equal cardinality (S == T)
  int64_t diag = INT64_MIN;          // skip pairs with i - j == diag
  uint64_t s_chunk = 0;              // stream sets per workgroup row (0: auto)
  uint32_t threads = 1024;           // workgroup size: 256, 512 or 1024
  int mask_bits = 0;                 // tile sets per lane: 64 (when the tile fits LDS), 32, or 0 (widest)
  uint32_t target_wgs = 1024;        // workgroups aimed at when s_chunk is auto
  unsigned long long* count_out = nullptr;  // COUNT: += number of pairs
  int* found = nullptr;                     // ANY: set to 1 on a hit
  unsigned long long* witness = nullptr;    // ANY: min of (i << 32) | j
  void* flags = nullptr;  // FLAGS / SFLAGS: bitmap over tile / stream sets (bit j set on a hit)
  uint32_t* scount = nullptr;  // SCOUNT: += pairs per stream set
  // Stream index of the indexed kernel (k_vsub_ix): per pair i of stream
  // sets, the LDS row offsets of both sets' states, interleaved by four, in
  // ix_list[ix_start[i] .. ix_start[i + 1]); |S_s| in ix_size[s].
  const uint32_t* ix_start = nullptr;
  const uint32_t* ix_list = nullptr;
  const uint32_t* ix_size = nullptr;
  const uint32_t* ix_head = nullptr;  // with ix_early_exit: each pair's first groups, fixed stride
  // Indexed kernel: list each stream set's states rarest-in-tile first
  // (fewest tile sets hold them), so that with ix_early_exit a stream set that
  // lies in no tile set is dropped after few states.
  int order_states = 0;
  // Indexed kernel, without diag: sort both lists by
  // cardinality first (a stream set can only lie in a tile set at least as
  // large), so each workgroup stops at the first stream set larger than its
  // largest tile set, and tiles hold sets of like size.
  int sort_by_size = 0;
  // The lists hold no equal sets other than the pairs diag names (both are
  // one deduplicated list): the indexed kernel may then skip equal sets by
  // cardinality (proper_by_size) instead of by diag, which sorting allows.
  int distinct_sets = 0;
  // Set by run_subset when sorted: sorted position -> list index of each
  // list, and s_bound[c] = stream sets of at most c states.
  const uint32_t* s_perm = nullptr;
  const uint32_t* t_perm = nullptr;
  const uint32_t* s_bound = nullptr;
  const uint32_t* t_bound = nullptr;  // t_bound[c] = tile sets of at most c states
  // Indexed kernel, ANY/FLAGS/SFLAGS without diag or proper_by_size: test
  // each stream set only against the tile sets holding its anchor (its state
  // held by the fewest tile sets), which a subset of it must hold.
  int bucket_anchor = 0;
  // Set by run_subset when bucketed: one workgroup per item, x = first entry
  // of t_perm (the anchor's inverted list), y = tile sets (<= 4096), [z, w) =
  // stream sets (z even) in s_perm order.
  const uint4* work = nullptr;
  uint64_t work_count = 0;
};

// Indexed subset kernel (k_subset_ix.hip): whether its tile fits LDS for n
// states, the pre-pass kernels that build the stream index, and the launch.
// The pre-pass kernels read set s of the list as sets[perm[s]] when perm is
// given (a list sorted by size), else sets[s].
bool vsub_ix_fits(uint32_t n, int W, bool proper_by_size);
hipError_t launch_ix_pair_count(int W, const uint64_t* sets, const uint32_t* perm, uint64_t count, uint32_t n,
    int complement, uint32_t* size, uint32_t* entries, hipStream_t st);
hipError_t launch_ix_pair_fill(int W, const uint64_t* sets, const uint32_t* perm, uint64_t count, uint32_t n,
    int complement, const uint32_t* start, const uint32_t* rank, const uint32_t* order, uint32_t* list,
    uint32_t* head, hipStream_t st);
// Groups of four states per half kept at a fixed stride per pair (ix_head):
// ix_head holds pairs x ix_head_groups() x 32 bytes.
int ix_head_groups();
// Sort by cardinality (counting sort; order within one cardinality is
// arbitrary): from hist (launch_size_hist), cursor[c] = sets of fewer than c
// states and, when bound != NULL, bound[c] = sets of at most c states; then
// perm[cursor[|S_i|]++] = i.
hipError_t launch_size_offsets(const uint32_t* hist, uint32_t n, uint32_t* cursor, uint32_t* bound, hipStream_t st);
hipError_t launch_size_scatter(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement,
    uint32_t* cursor, uint32_t* perm, hipStream_t st);
// Per state q < n, freq[q] += number of sets (complemented when `complement`)
// holding q; freq must start at zero.
hipError_t launch_state_freq(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement, uint32_t* freq,
    hipStream_t st);
// hist[c] += number of sets (complemented when `complement`) of c states,
// 0 <= c <= n; hist must start at zero.  W <= 5.
hipError_t launch_size_hist(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement, uint32_t* hist,
    hipStream_t st);
// Anchor buckets: anchor[i] = the stream set's state of lowest rank (n if
// empty), hist[anchor] += 1; perm[cursor[key[i]]++] = i for keys < nkeys;
// lists[cursor[q]++] = t for every state q of every tile set t.
hipError_t launch_anchor(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement,
    const uint32_t* rank, uint32_t* anchor, uint32_t* hist, hipStream_t st);
hipError_t launch_key_scatter(const uint32_t* key, uint64_t count, uint32_t nkeys, uint32_t* cursor, uint32_t* perm,
    hipStream_t st);
hipError_t launch_invert(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement, uint32_t* cursor,
    uint32_t* lists, hipStream_t st);
// rank[q] = position of state q in ascending (freq, q) order; order[r] = the
// LDS row byte offset of the state at position r.
hipError_t launch_state_rank(const uint32_t* freq, uint32_t n, uint32_t* rank, uint32_t* order, hipStream_t st);

// Device subset trie (k_trie.hip; built by trie.cpp).  A node is a uint4:
// internal {zero child, one child, bit | minpop << 16, maxpop}, leaf {first
// set, set count, kTrieLeafBit | minpop << 16, maxpop}; node 0 is the root,
// and no child is node 0.
constexpr uint32_t kTrieLeafBit = 0xffffu;
constexpr uint32_t kTrieMaxDepth = 48;  // nodes at this depth are leaves
constexpr uint32_t kTrieLeafSets = 8;   // leaves hold at most this many sets above the depth cap
constexpr uint32_t kTrieChunk = 2048;   // sets per workgroup of the level-synchronous passes
// Largest node built whole in one workgroup's LDS (two set and two tag buffers).
__host__ __device__ constexpr uint32_t trie_local_cap(int W) {
  return W == 1 ? 4096u : (W <= 3 ? 2048u : (W <= 7 ? 1024u : 512u));
}
__host__ __device__ constexpr size_t trie_local_lds(int W) {
  return 2 * size_t(trie_local_cap(W)) * (8 * size_t(W) + 4);
}
struct TrieSeg {
  uint32_t node, begin, end, depth;
};
struct TrieChunk {
  uint32_t seg, begin, end, pad;
};
// A node built by one workgroup (k_trie_local; k_trie_bigleaf for a leaf
// too large for LDS): rows [begin, end) of buffer buf (0 input, 1, 2).
struct TrieLocalJob {
  uint32_t root, begin, end, depth, buf, desc_base, pad0, pad1;
};
struct TrieBuildArgs {
  const uint64_t* in;        // input sets (complemented on load when in_cmask is all ones)
  uint64_t in_cmask;
  const uint32_t* in_tags;   // tags of the input sets (NULL: row index)
  const uint64_t* buf1;
  const uint64_t* buf2;
  const uint32_t* tags1;
  const uint32_t* tags2;
  uint64_t* out;             // the trie's sets, in trie order
  uint32_t* out_tags;
  uint4* nodes;
  uint32_t n;
};
struct TrieQueryArgs {
  const uint4* nodes;
  uint32_t node_count;
  const uint64_t* tsets;
  const uint32_t* ttags;
  const uint64_t* queries;
  uint64_t q_cmask;          // all ones: the queries are stored complemented
  uint32_t qcount, n;
  const uint32_t* perm;      // query order (NULL: list order)
  int reverse;               // walk perm from its end
  int sorted;                // perm is a cardinality order in the walk's favour
  int query_high;            // ANY witness: (query << 32) | tag, else (tag << 32) | query
  uint32_t* next;            // work counter, zero at launch
  int* found;
  unsigned long long* witness;
  uint8_t* flags;            // FLAGS: flags[query] = 1 on a hit
};
hipError_t launch_trie_hist(int W, const uint64_t* src, int complement, uint32_t n, const TrieChunk* chunks,
    uint32_t nchunks, uint32_t* hist, uint32_t* seg_min, uint32_t* seg_max, hipStream_t st);
hipError_t launch_trie_choose(const uint32_t* hist, const TrieSeg* segs, uint32_t nsegs, uint32_t n, uint32_t* bits,
    uint32_t* ones, hipStream_t st);
hipError_t launch_trie_chunk_zeros(int W, const uint64_t* src, int complement, const TrieChunk* chunks,
    uint32_t nchunks, const uint32_t* bits, uint32_t* czero, hipStream_t st);
hipError_t launch_trie_scatter(int W, const uint64_t* src, int complement, uint32_t n, const uint32_t* src_tags,
    uint64_t* dst, uint32_t* dst_tags, const TrieChunk* chunks, uint32_t nchunks, const uint32_t* bits,
    const uint2* coff, hipStream_t st);
hipError_t launch_trie_bigleaf(int W, const TrieBuildArgs& a, const TrieLocalJob* jobs, uint32_t njobs,
    hipStream_t st);
hipError_t launch_trie_local(int W, const TrieBuildArgs& a, const TrieLocalJob* jobs, uint32_t njobs,
    hipStream_t st);
// superset: trie sets containing the query (else inside it); any: stop at
// the first hit with a witness (else per-query flags).
hipError_t launch_trie_query(int W, int superset, int any, const TrieQueryArgs& a, uint32_t grid, hipStream_t st);
// Cardinality order (any W): hist[c] += sets of c states; then with
// launch_size_offsets' cursor, perm[cursor[|S_i|]++] = i.
hipError_t launch_card_hist(int W, const uint64_t* sets, uint64_t count, int complement, uint32_t n, uint32_t* hist,
    hipStream_t st);
hipError_t launch_card_scatter(int W, const uint64_t* sets, uint64_t count, int complement, uint32_t n,
    uint32_t* cursor, uint32_t* perm, hipStream_t st);

// Beam of an automaton of at most 64 states in one workgroup
// (k_beam_small.hip): every depth in one launch while a depth's candidates
// fit kBeamSmallCand and the sets it keeps kBeamSmallList; out[0] = status, out[1] = the depth it stopped at (the
// bound when found), out[2] = the list size there (list_out), out[3] = the
// history entries written; meta[2 d], meta[2 d + 1] = history offset and list
// size of depth d + 1; word[0 .. bound) on a find.
constexpr uint32_t kBeamSmallCand = 4096;  // candidates of one depth
constexpr uint32_t kBeamSmallList = 2048;  // sets a depth keeps
constexpr uint32_t kBeamSmallMaxK = 16;
enum BeamSmallStatus : uint32_t {
  kBeamSmallFound = 1,      // a depth produced Q
  kBeamSmallExhausted = 2,  // every depth below upper (or the stop cell's depth) done, no Q
  kBeamSmallEmpty = 3,      // the list became empty
  kBeamSmallHandover = 4,   // a depth too large for LDS: continue on the device
};
struct BeamSmallArgs {
  uint32_t n, k, width, upper, m0;
  const uint64_t* seed;
  const uint32_t* delta;
  uint32_t* hist;
  uint32_t hist_cap;
  uint32_t* meta;
  uint32_t meta_cap;
  uint32_t* word;
  uint64_t* list_out;
  uint32_t* out;
  // Optional stop cell in pinned host memory, (generation << 32) | depth,
  // written by the host while the beam runs: every 16 depths the beam reads
  // it and ends (no Q) once it holds stop_gen and len >= depth.
  const unsigned long long* stop;
  uint32_t stop_gen;
};
hipError_t launch_beam_small(const BeamSmallArgs& a, hipStream_t st);

// Power-set BFS (k_brute.hip).  One automaton per workgroup with its bitmap
// in LDS (brute_lds_bytes(n, k) <= kBruteMaxLds): frontier_off
// indexes two buffers of 2^n subsets each.  Or one level over the device
// (k <= kBruteLevelMaxK): state[0] += new subsets, state[1] = 1 on a
// singleton.
struct BruteJob {
  uint32_t n, k;
  uint64_t delta_off;     // into deltas, n * k entries
  uint64_t frontier_off;  // into frontier, 2 * 2^n entries
};
constexpr int kBruteLevelMaxK = 8;
constexpr size_t kBruteMaxLds = 160 * 1024 - 64;  // dynamic LDS of k_brute_wg (64 bytes static)
int brute_lds_bytes(uint32_t n, uint32_t k);
hipError_t launch_brute_wg(const BruteJob* jobs, uint32_t count, const uint32_t* deltas, uint32_t* frontier,
    int64_t* out, size_t lds, hipStream_t st);
hipError_t launch_brute_level(uint32_t n, uint32_t k, const uint32_t* delta, const uint32_t* cur, const uint32_t* fc,
    uint32_t fc_host, uint32_t* nxt, uint32_t* map, uint32_t* state, hipStream_t st);
hipError_t launch_vsub_ix(int W, const VsubParams& p, hipStream_t st);

// Eppstein's greedy bound, one workgroup per automaton (k_eppstein.hip);
// n <= kEppMaxN.  Scratch per automaton: eppstein_scratch_bytes(n, k) at
// scratch_off; the word (up to word_cap letters) at words + word_off.
constexpr uint32_t kEppMaxN = 512;
struct EppJob {
  uint32_t n, k;
  uint64_t delta_off, scratch_off, word_off, word_cap, upper;
};
struct EppOut {
  int32_t non_synchro, improved;  // improved: the greedy length is at most upper
  uint64_t bound, word_len;       // bound = upper unless improved
};
// Scratch layout of one automaton, every region 8-byte aligned (keys take
// 64-bit atomics, which must be naturally aligned): beg[k n + 1] (u32),
// edges[k n] (u16), info[n n] (u32), keys[n n] (u64), queue and per-parent
// counts (n (n - 1) / 2 + k n u32 each).
struct EppLayout {
  size_t edges, info, keys, queue, cnt, total;
};
__host__ __device__ inline EppLayout epp_layout(uint32_t n, uint32_t k) {
  const size_t kn = size_t(k) * n, nn = size_t(n) * n, all = size_t(n) * (n - 1) / 2;
  auto a8 = [](size_t b) { return (b + 7) & ~size_t(7); };
  EppLayout l;
  l.edges = a8((kn + 1) * 4);
  l.info = l.edges + a8(kn * 2);
  l.keys = l.info + a8(nn * 4);
  l.queue = l.keys + nn * 8;
  l.cnt = l.queue + a8((all + kn) * 4);
  l.total = (l.cnt + (all + kn) * 4 + 255) & ~size_t(255);
  return l;
}
size_t eppstein_scratch_bytes(uint32_t n, uint32_t k);
// LDS bytes of an automaton's preimage lists, 0 when they are too large to
// be staged there (the kernels then read them from the scratch).
uint32_t eppstein_list_lds(uint32_t n, uint32_t k);
// Bytes of the greedy walk's pair letters (max_n * max_n) when they fit the
// launch's dynamic LDS after list_lds bytes of lists, else 0 (the walk then
// reads them from the scratch).
uint32_t eppstein_letter_lds(uint32_t max_n, uint32_t list_lds);
// list_lds: the dynamic LDS of the launch, the largest eppstein_list_lds of
// its jobs (a job whose lists need more reads them from the scratch);
// letter_lds: eppstein_letter_lds of its largest n; small: every job has at
// most 64 states (the pair tables then live in LDS, about 56 KB per
// workgroup, which larger launches do without to keep two per CU).
hipError_t launch_eppstein_wg(const EppJob* jobs, uint32_t count, const uint32_t* deltas, unsigned char* scratch,
    uint32_t* words, EppOut* outs, uint32_t list_lds, uint32_t letter_lds, bool small, hipStream_t st);
// The pair BFS of one automaton over the whole device.  EppLevel is its
// device state: the level's parents at queue[head, tail) (level 0: the k n
// preimage-list pairs), pairs found so far, done, the last level's winners.
// launch_epp_level enqueues one level (base = the automaton's scratch,
// epp_layout, with beg and edges filled; off: n (n - 1) / 2 + k n words;
// tile_sum: epp_scan_tiles words) without waiting; levels enqueued after the
// last one return at once.  Then the greedy phase in one workgroup (job: one
// EppJob on the device).
struct EppLevel {
  uint32_t head, tail, found, done, total;
};
hipError_t launch_epp_level(uint32_t level, uint32_t n, uint32_t k, unsigned char* base, uint32_t* off,
    uint32_t* tile_sum, EppLevel* st, hipStream_t stream);
uint32_t epp_scan_tiles(uint32_t n, uint32_t k);
hipError_t launch_epp_greedy(const EppJob* job, uint32_t n, const uint32_t* deltas, const unsigned char* scratch,
    uint32_t* words, EppOut* out, hipStream_t st);

// Mask width and LDS bytes of the subset kernel.
int vsub_mask_bits(uint32_t n, int W, int requested, bool proper_by_size);
size_t vsub_lds_bytes(uint32_t n, int W, int mask_bits, bool proper_by_size);
// Bytes of the FLAGS bitmap for t tile sets (whole 64-bit words; bit j of
// the bitmap, read as bytes, is tile set j for either mask width).
inline size_t vsub_flag_bytes(int W, uint64_t t) {
  (void)W;
  return size_t((t + 63) / 64) * 8;
}
hipError_t launch_vsub(int W, const VsubParams& p, hipStream_t st);
// Bernoulli(p) random sets (counter-based hash of seed and position).
hipError_t launch_random_sets(int W, uint64_t* out, uint64_t count, uint32_t n, float p, uint64_t seed,
    hipStream_t st);

// out[a*m + i] = image of parents[i] under letter a, from 8-state slice
// tables tables[a][s][256][W] (the reference's PreprocessedTransition
// layout).  valid[a*m + i] = 0 when the child is to be dropped:
// drop_mode 1 drops children of >= n-1 states (complemented sets of at most
// one state).  valid may be NULL (nothing dropped).
hipError_t launch_expand(int W, const uint64_t* parents, uint64_t m, const uint64_t* tables, uint32_t slices,
    uint32_t letters, const uint32_t* letter_list, uint32_t n, int drop_mode, uint64_t* out, uint32_t* valid,
    hipStream_t st);
// The same, for launches past 2^17 children bit-sliced (64 parents per wave,
// transposed across the lanes; targets[a * n + q] = delta(q, a), u16);
// smaller launches, or targets NULL, take the slice tables.
hipError_t launch_expand_sets(int W, const uint64_t* parents, uint64_t m, const uint64_t* tables, uint32_t slices,
    const uint16_t* targets, int dir, uint32_t letters, const uint32_t* letter_list, uint32_t n, int drop_mode,
    uint64_t* out, uint32_t* valid, hipStream_t st);

// Open-addressing hash set over rows of an arena.  Slots hold an arena index,
// or kCand | candidate index while a step is being merged.
constexpr uint32_t kEmpty = 0xffffffffu;
constexpr uint32_t kCand = 0x80000000u;
hipError_t launch_hash_insert(int W, const uint64_t* cand, const uint32_t* valid, uint64_t count, uint32_t* table,
    uint64_t mask, const uint64_t* arena, uint32_t* slot_of, unsigned long long* probes, hipStream_t st);
// keep[i] = 1 iff candidate i owns its slot (smallest index among equal sets
// that are new).
hipError_t launch_hash_resolve(const uint32_t* slot_of, const uint32_t* table, uint64_t count, uint32_t* keep,
    hipStream_t st);
// Exclusive prefix sum of keep into pos; total into *total (device).
hipError_t launch_scan(const uint32_t* keep, uint32_t* pos, uint64_t count, uint32_t* total, void* temp,
    size_t* temp_bytes, hipStream_t st);
// Append kept candidates to the arena at `base` (+pos[i]); record parent
// (parent_idx[i % m]) and letter (letter_list[i / m]); point the hash slot
// at the new row.
// (parent_idx NULL: parent = parent_base + i % m.)
// front != NULL: the kept candidates are also written at front[pos] (with
// their rows at front_idx[pos]); found != NULL: *found = 0 and *witness = ~0
// (as launch_goal_reset) in the same launch.
hipError_t launch_commit(int W, const uint64_t* cand, const uint32_t* keep, const uint32_t* pos,
    const uint32_t* slot_of, uint64_t count, uint64_t m, const uint32_t* parent_idx, uint64_t parent_base,
    const uint32_t* letter_list, uint64_t base, uint64_t* arena, uint32_t* parent, uint8_t* letter, uint32_t* table,
    hipStream_t st, uint64_t* front = nullptr, uint32_t* front_idx = nullptr, int* found = nullptr,
    unsigned long long* witness = nullptr);
// Insert arena rows [0, rows) (all distinct) into an empty table.
// valid[i] = 0 for each valid candidate equal to an arena row in the table.
hipError_t launch_hash_lookup(int W, const uint64_t* cand, uint32_t* valid, uint64_t count, const uint32_t* table,
    uint64_t mask, const uint64_t* arena, hipStream_t st);
hipError_t launch_hash_rebuild(int W, const uint64_t* arena, uint64_t rows, uint32_t* table, uint64_t mask,
    hipStream_t st);
// Slice tables [k][ceil(n/8)][256][W] from per-state masks [k][n][W] (row
// (a, s, b) = OR of the masks of the states 8 s + j, j a set bit of b).
hipError_t launch_tables(int W, const uint64_t* masks, uint32_t n, uint32_t k, uint64_t* out, hipStream_t st);
// keep[i] = !(bit i of dead) for i < count.
hipError_t launch_alive(int W, const void* dead_bits, uint64_t count, uint32_t* keep, hipStream_t st);
// Keep flags (mode 0 given, 1 = hash-slot owners as launch_hash_resolve,
// 2 = clear bits of dead_bits as launch_alive) and their exclusive scan with
// the total: one launch up to 2048 entries, two up to 2^21, the scan above
// beyond.  temp holds keep_scan_temp_bytes(count) bytes.
hipError_t keep_scan_temp_bytes(uint64_t count, size_t* temp_bytes);
hipError_t launch_keep_scan(int mode, const uint32_t* slot_of, const uint32_t* table, const void* dead_bits,
    uint64_t count, uint32_t* keep, uint32_t* pos, uint32_t* total, void* temp, size_t* temp_bytes, hipStream_t st);
// *found = 0, *witness = ~0 (before a meet check).
hipError_t launch_goal_reset(int* found, unsigned long long* witness, hipStream_t st);
// out[j] = rows[lo + i] for kept i (j = pos[i]); idx_out[j] = lo + i.
hipError_t launch_select(int W, const uint64_t* rows, uint64_t lo, const uint32_t* keep, const uint32_t* pos,
    uint64_t count, uint64_t* out, uint32_t* idx_out, hipStream_t st);
// The device beam's compaction of one depth's distinct candidates: as
// launch_select with lo = 0, and idx_copy[j] = i too; *first = i for the kept
// candidate that is Q (all n states), if any (the smallest index of the full
// candidates, since the kept one of equal candidates is the smallest); and
// table[slot_of[i]] = empty for every kept i, which empties every slot the
// depth's hash insertion filled (each holds one kept candidate).
hipError_t launch_beam_select(int W, const uint64_t* cand, const uint32_t* keep, const uint32_t* pos,
    uint64_t count, uint32_t n, uint64_t* out, uint32_t* idx_out, uint32_t* idx_copy, unsigned long long* first,
    uint32_t* table, const uint32_t* slot_of, hipStream_t st);
// Per-state frequency over a list (acc must be zeroed): acc[q] += #sets with q.
hipError_t launch_count_bits(int W, const uint64_t* sets, uint64_t count, uint32_t n, unsigned long long* acc,
    hipStream_t st);
// Streaming copy of bytes/16 16-byte vectors (traffic calibration).
hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t st);
// LDS read calibration: `blocks` workgroups of 1024 threads, each wave reading
// iters * 8 rows of 64 lanes x bytes_per_lane (8 or 16) bytes.
hipError_t launch_lds_read(int bytes_per_lane, uint64_t iters, uint32_t blocks, uint32_t* out, hipStream_t st);
// Fill 32-bit words.
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t count, hipStream_t st);

// Reduce stage (k_reduce.hip).  out[w] |= the OR of word w over `count` rows
// of W words (out zeroed by the caller).
hipError_t launch_or_rows(const uint64_t* rows, uint64_t count, uint32_t W, unsigned long long* out, hipStream_t st);
// *first = min(*first, the lowest index of a row with exactly one state).
hipError_t launch_first_singleton(const uint64_t* rows, uint64_t count, uint32_t W, unsigned long long* first,
    hipStream_t st);
// out row i (w_out words) = { map[q] : q in in row i (w_in words) }; map has
// n_in entries, defined for every state the rows hold.
hipError_t launch_remap_sets(const uint64_t* in, uint64_t count, uint32_t w_in, const uint16_t* map, uint32_t n_in,
    uint64_t* out, uint32_t w_out, hipStream_t st);

}  // namespace sw
