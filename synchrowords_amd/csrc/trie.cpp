// Host driver of the device subset trie (k_trie.hip) and its C-ABI seam.
//
// The reference builds its SubsetsTrie recursively on the host
// (synchrolib/data_structures/subsets_trie.hpp:57-70, 236-297): split a
// node's sets on the most frequent non-constant state, recurse, stop at
// M = 10 sets.  Here the top of the trie, where nodes hold more sets than
// one workgroup's LDS, is built one level at a time over the whole device
// (one read-back per level: each node's division bit, its set count with
// the bit, and per 2048-set chunk the sets without it, from which the host
// places every chunk's two halves); every node small enough is handed whole
// to one workgroup (k_trie_local).  Node ids: the host numbers the top nodes
// and the roots of the workgroup-built subtrees; each of those subtrees
// numbers its own nodes inside a range of 2 x its set count reserved after
// all top ids.
#include "sw_trie.hpp"

#include <algorithm>
#include <cstring>
#include <memory>

namespace sw {

void DeviceTrie::build(hipStream_t st, uint32_t n_, const uint64_t* src, uint64_t count, int complement,
    const uint32_t* tags_in) {
  n = n_;
  W = (n + 63) / 64;
  m = count;
  node_count = 0;
  top_levels = 0;
  if (!m) return;
  SW_REQUIRE(m < (1ull << 31), "trie: more than 2^31 sets");
  const uint32_t cap = trie_local_cap(int(W));
  const size_t row = size_t(W) * 8;
  sets.ensure(m * row);
  tags.ensure(m * 4);
  TrieBuildArgs a{};
  a.in = src;
  a.in_cmask = complement ? ~0ull : 0ull;
  a.in_tags = tags_in;
  a.out = sets.as<uint64_t>();
  a.out_tags = tags.as<uint32_t>();
  a.n = n;
  std::vector<TrieLocalJob> local, big;
  std::vector<uint4> top(1, make_uint4(0, 0, 0, 0));  // host-numbered node records (placeholders for jobs)
  uint32_t next_id = 1;
  if (m <= cap) {
    local.push_back(TrieLocalJob{0, 0, uint32_t(m), 0, 0, 0, 0, 0});
  } else {
    buf1_.ensure(m * row);
    buf2_.ensure(m * row);
    tags1_.ensure(m * 4);
    tags2_.ensure(m * 4);
    a.buf1 = buf1_.as<uint64_t>();
    a.buf2 = buf2_.as<uint64_t>();
    a.tags1 = tags1_.as<uint32_t>();
    a.tags2 = tags2_.as<uint32_t>();
    const uint64_t* bufs[3] = {src, a.buf1, a.buf2};
    const uint32_t* tbufs[3] = {tags_in, a.tags1, a.tags2};
    uint64_t* wbufs[3] = {nullptr, buf1_.as<uint64_t>(), buf2_.as<uint64_t>()};
    uint32_t* wtbufs[3] = {nullptr, tags1_.as<uint32_t>(), tags2_.as<uint32_t>()};
    std::vector<TrieSeg> segs{TrieSeg{0, 0, uint32_t(m), 0}};
    int cur = 0;
    while (!segs.empty()) {
      ++top_levels;
      const int nxt = cur == 1 ? 2 : 1;
      const uint32_t S = uint32_t(segs.size());
      std::vector<TrieChunk> chunks;
      std::vector<uint32_t> first_chunk(S + 1);
      for (uint32_t s = 0; s < S; ++s) {
        first_chunk[s] = uint32_t(chunks.size());
        for (uint32_t b = segs[s].begin; b < segs[s].end; b += kTrieChunk)
          chunks.push_back(TrieChunk{s, b, std::min(b + kTrieChunk, segs[s].end), 0});
      }
      first_chunk[S] = uint32_t(chunks.size());
      const uint32_t C = uint32_t(chunks.size());
      TrieSeg* hs = h_segs_.get<TrieSeg>(S);
      std::memcpy(hs, segs.data(), S * sizeof(TrieSeg));
      TrieChunk* hc = h_chunks_.get<TrieChunk>(C);
      std::memcpy(hc, chunks.data(), C * sizeof(TrieChunk));
      segs_.ensure(S * sizeof(TrieSeg));
      chunks_.ensure(C * sizeof(TrieChunk));
      hist_.ensure(size_t(S) * n * 4);
      segmin_.ensure(S * 4);
      segmax_.ensure(S * 4);
      bits_.ensure(S * 4);
      ones_.ensure(S * 4);
      czero_.ensure(C * 4);
      coff_.ensure(C * 8);
      SW_CHECK_HIP(hipMemcpyAsync(segs_.p, hs, S * sizeof(TrieSeg), hipMemcpyHostToDevice, st));
      SW_CHECK_HIP(hipMemcpyAsync(chunks_.p, hc, C * sizeof(TrieChunk), hipMemcpyHostToDevice, st));
      SW_CHECK_HIP(launch_fill_u32(hist_.as<uint32_t>(), 0u, uint64_t(S) * n, st));
      SW_CHECK_HIP(launch_fill_u32(segmin_.as<uint32_t>(), 0xffffffffu, S, st));
      SW_CHECK_HIP(launch_fill_u32(segmax_.as<uint32_t>(), 0u, S, st));
      const int cm = cur == 0 ? complement : 0;
      SW_CHECK_HIP(launch_trie_hist(int(W), bufs[cur], cm, n, chunks_.as<TrieChunk>(), C, hist_.as<uint32_t>(),
          segmin_.as<uint32_t>(), segmax_.as<uint32_t>(), st));
      SW_CHECK_HIP(launch_trie_choose(hist_.as<uint32_t>(), segs_.as<TrieSeg>(), S, n, bits_.as<uint32_t>(),
          ones_.as<uint32_t>(), st));
      SW_CHECK_HIP(launch_trie_chunk_zeros(int(W), bufs[cur], cm, chunks_.as<TrieChunk>(), C, bits_.as<uint32_t>(),
          czero_.as<uint32_t>(), st));
      uint32_t* back = h_back_.get<uint32_t>(4 * size_t(S) + C);
      SW_CHECK_HIP(hipMemcpyAsync(back, bits_.p, S * 4, hipMemcpyDeviceToHost, st));
      SW_CHECK_HIP(hipMemcpyAsync(back + S, ones_.p, S * 4, hipMemcpyDeviceToHost, st));
      SW_CHECK_HIP(hipMemcpyAsync(back + 2 * S, segmin_.p, S * 4, hipMemcpyDeviceToHost, st));
      SW_CHECK_HIP(hipMemcpyAsync(back + 3 * S, segmax_.p, S * 4, hipMemcpyDeviceToHost, st));
      SW_CHECK_HIP(hipMemcpyAsync(back + 4 * S, czero_.p, C * 4, hipMemcpyDeviceToHost, st));
      SW_CHECK_HIP(hipStreamSynchronize(st));
      const uint32_t* bits = back;
      const uint32_t* ones = back + S;
      const uint32_t* smin = back + 2 * S;
      const uint32_t* smax = back + 3 * S;
      const uint32_t* cz = back + 4 * S;
      uint2* hco = h_coff_.get<uint2>(C);
      std::vector<TrieSeg> next;
      bool any_split = false;
      for (uint32_t s = 0; s < S; ++s) {
        const TrieSeg g = segs[s];
        if (bits[s] == kTrieLeafBit) {
          // Every state constant over more sets than LDS takes (equal sets):
          // one leaf, copied by one workgroup.
          big.push_back(TrieLocalJob{g.node, g.begin, g.end, g.depth, uint32_t(cur), 0, 0, 0});
          continue;
        }
        any_split = true;
        const uint32_t zc = (g.end - g.begin) - ones[s];
        const uint32_t z = next_id++, o = next_id++;
        top.resize(next_id, make_uint4(0, 0, 0, 0));
        top[g.node] = make_uint4(z, o, bits[s] | (smin[s] << 16), smax[s]);
        const TrieSeg kids[2] = {TrieSeg{z, g.begin, g.begin + zc, g.depth + 1},
                                 TrieSeg{o, g.begin + zc, g.end, g.depth + 1}};
        for (const TrieSeg& kd : kids) {
          const uint32_t size = kd.end - kd.begin;
          if (size <= cap) local.push_back(TrieLocalJob{kd.node, kd.begin, kd.end, kd.depth, uint32_t(nxt), 0, 0, 0});
          else if (kd.depth >= kTrieMaxDepth) big.push_back(TrieLocalJob{kd.node, kd.begin, kd.end, kd.depth, uint32_t(nxt), 0, 0, 0});
          else next.push_back(kd);
        }
        uint32_t zb = g.begin, ob = g.begin + zc;
        for (uint32_t c = first_chunk[s]; c < first_chunk[s + 1]; ++c) {
          hco[c] = make_uint2(zb, ob);
          zb += cz[c];
          ob += (chunks[c].end - chunks[c].begin) - cz[c];
        }
      }
      if (any_split) {
        SW_CHECK_HIP(hipMemcpyAsync(coff_.p, hco, C * 8, hipMemcpyHostToDevice, st));
        SW_CHECK_HIP(launch_trie_scatter(int(W), bufs[cur], cm, n, tbufs[cur], wbufs[nxt], wtbufs[nxt],
            chunks_.as<TrieChunk>(), C, bits_.as<uint32_t>(), coff_.as<uint2>(), st));
      }
      segs.swap(next);
      cur = nxt;
    }
  }
  // Id ranges of the workgroup-built subtrees, after every host-numbered id.
  for (TrieLocalJob& j : local) {
    j.desc_base = next_id;
    next_id += 2 * (j.end - j.begin);
  }
  node_count = next_id;
  nodes.ensure(size_t(node_count) * sizeof(uint4));
  const size_t ntop = top.size();
  uint4* hn = h_nodes_.get<uint4>(ntop);
  std::memcpy(hn, top.data(), ntop * sizeof(uint4));
  SW_CHECK_HIP(hipMemcpyAsync(nodes.p, hn, ntop * sizeof(uint4), hipMemcpyHostToDevice, st));
  a.nodes = nodes.as<uint4>();
  const size_t nj = local.size() + big.size();
  TrieLocalJob* hj = h_jobs_.get<TrieLocalJob>(nj);
  std::copy(local.begin(), local.end(), hj);
  std::copy(big.begin(), big.end(), hj + local.size());
  jobs_.ensure(nj * sizeof(TrieLocalJob));
  SW_CHECK_HIP(hipMemcpyAsync(jobs_.p, hj, nj * sizeof(TrieLocalJob), hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(launch_trie_local(int(W), a, jobs_.as<TrieLocalJob>(), uint32_t(local.size()), st));
  SW_CHECK_HIP(launch_trie_bigleaf(int(W), a, jobs_.as<TrieLocalJob>() + local.size(), uint32_t(big.size()), st));
  // The pinned job and node records must outlive their copies.
  SW_CHECK_HIP(hipStreamSynchronize(st));
}

void DeviceTrie::query(hipStream_t st, uint32_t grid, int superset, const uint64_t* q, uint64_t qcount,
    int q_complement, int sort, int query_high, int* found, unsigned long long* witness, uint8_t* flags) {
  if (!m || !qcount) return;
  SW_REQUIRE(qcount < (1ull << 32), "trie query: more than 2^32 queries");
  TrieQueryArgs a{};
  a.nodes = nodes.as<uint4>();
  a.node_count = node_count;
  a.tsets = sets.as<uint64_t>();
  a.ttags = tags.as<uint32_t>();
  a.queries = q;
  a.q_cmask = q_complement ? ~0ull : 0ull;
  a.qcount = uint32_t(qcount);
  a.n = n;
  qnext_.ensure(64);
  SW_CHECK_HIP(launch_fill_u32(qnext_.as<uint32_t>(), 0u, 1, st));
  a.next = qnext_.as<uint32_t>();
  a.found = found;
  a.witness = witness;
  a.flags = flags;
  a.query_high = query_high;
  if (sort) {
    // Cardinality order (the reference queries its trie one cardinality at
    // a time, dfs.cpp:263-344): largest first for subset queries, smallest
    // first for superset ones, so that every batch after the first one
    // pruned at the root is pruned too.
    qhist_.ensure(size_t(n + 1) * 4);
    qcursor_.ensure(size_t(n + 1) * 4);
    qperm_.ensure(qcount * 4);
    SW_CHECK_HIP(launch_fill_u32(qhist_.as<uint32_t>(), 0u, n + 1, st));
    SW_CHECK_HIP(launch_card_hist(int(W), q, qcount, q_complement, n, qhist_.as<uint32_t>(), st));
    SW_CHECK_HIP(launch_size_offsets(qhist_.as<uint32_t>(), n, qcursor_.as<uint32_t>(), nullptr, st));
    SW_CHECK_HIP(launch_card_scatter(int(W), q, qcount, q_complement, n, qcursor_.as<uint32_t>(),
        qperm_.as<uint32_t>(), st));
    a.perm = qperm_.as<uint32_t>();
    a.reverse = superset ? 0 : 1;
    a.sorted = 1;
  }
  SW_CHECK_HIP(launch_trie_query(int(W), superset, found ? 1 : 0, a, grid, st));
}

}  // namespace sw

// ------------------------------------------------------------------ C ABI

struct sw_trie {
  sw_ctx* ctx = nullptr;
  sw::DeviceTrie trie;
  sw::DevBuf in, tags, q, out, small;
};

namespace {
uint32_t trie_grid(sw_ctx* ctx) { return uint32_t(std::max(1, ctx->cus)) * 3; }
}  // namespace

extern "C" int sw_trie_build(sw_ctx* ctx, uint32_t n, const uint64_t* sets, uint64_t count, int where,
    sw_trie** out) {
  try {
    SW_REQUIRE(ctx && out && (sets || !count), "NULL argument");
    SW_REQUIRE(n >= 1, "n must be >= 1");
    if (n > 512) throw sw::Error{SW_ERR_UNSUPPORTED, "n above 512 states is not compiled"};
    SW_REQUIRE(count < (1ull << 31), "more than 2^31 sets");
    sw::use(ctx);
    auto t = std::make_unique<sw_trie>();
    t->ctx = ctx;
    const uint32_t W = (n + 63) / 64;
    const uint64_t* src = sets;
    if (where == SW_HOST && count) {
      t->in.ensure(count * W * 8);
      SW_CHECK_HIP(hipMemcpyAsync(t->in.p, sets, count * W * 8, hipMemcpyHostToDevice, ctx->stream));
      src = t->in.as<uint64_t>();
    }
    t->trie.build(ctx->stream, n, src, count, 0, nullptr);
    SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    *out = t.release();
    return SW_OK;
  } catch (const sw::Error& e) {
    sw::set_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    sw::set_error(e.what());
    return SW_ERR_INVALID;
  }
}

extern "C" int sw_trie_query(sw_trie* t, int superset, const uint64_t* queries, uint64_t count, int where,
    uint8_t* results, int32_t* any, uint64_t* witness) {
  try {
    SW_REQUIRE(t && (queries || !count), "NULL argument");
    sw_ctx* ctx = t->ctx;
    sw::use(ctx);
    SW_REQUIRE(count < (1ull << 32), "more than 2^32 queries");
    const uint32_t W = (t->trie.n + 63) / 64;
    const uint64_t* q = queries;
    if (where == SW_HOST && count) {
      t->q.ensure(count * W * 8);
      SW_CHECK_HIP(hipMemcpyAsync(t->q.p, queries, count * W * 8, hipMemcpyHostToDevice, ctx->stream));
      q = t->q.as<uint64_t>();
    }
    t->small.ensure(64);
    int* dfound = t->small.as<int>();
    unsigned long long* dwit = reinterpret_cast<unsigned long long*>(t->small.as<char>() + 8);
    SW_CHECK_HIP(sw::launch_goal_reset(dfound, dwit, ctx->stream));
    if (results) {
      t->out.ensure(count + 8);
      SW_CHECK_HIP(hipMemsetAsync(t->out.p, 0, count + 8, ctx->stream));
      t->trie.query(ctx->stream, trie_grid(ctx), superset, q, count, 0, 1, 1, nullptr, nullptr,
          t->out.as<uint8_t>());
      if (count) SW_CHECK_HIP(hipMemcpyAsync(results, t->out.p, count, hipMemcpyDeviceToHost, ctx->stream));
      SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
      int32_t f = 0;
      for (uint64_t i = 0; i < count; ++i) f |= results[i];
      if (any) *any = f;
      if (witness) *witness = ~0ull;
    } else {
      t->trie.query(ctx->stream, trie_grid(ctx), superset, q, count, 0, 1, 1, dfound, dwit, nullptr);
      uint64_t host[2] = {0, 0};
      SW_CHECK_HIP(hipMemcpyAsync(host, t->small.p, 16, hipMemcpyDeviceToHost, ctx->stream));
      SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
      if (any) *any = int32_t(uint32_t(host[0]) != 0);
      if (witness) *witness = host[1];
    }
    return SW_OK;
  } catch (const sw::Error& e) {
    sw::set_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    sw::set_error(e.what());
    return SW_ERR_INVALID;
  }
}

extern "C" int sw_trie_info(sw_trie* t, uint64_t* sets, uint64_t* nodes, uint32_t* top_levels) {
  if (!t) return SW_ERR_INVALID;
  if (sets) *sets = t->trie.m;
  if (nodes) *nodes = t->trie.node_count;
  if (top_levels) *top_levels = t->trie.top_levels;
  return SW_OK;
}

extern "C" int sw_trie_nodes(sw_trie* t, uint32_t* out, uint64_t cap_nodes, uint64_t* out_sets, uint32_t* out_tags) {
  try {
    SW_REQUIRE(t, "NULL argument");
    sw_ctx* ctx = t->ctx;
    sw::use(ctx);
    const uint64_t nn = std::min<uint64_t>(cap_nodes, t->trie.node_count);
    if (out && nn) SW_CHECK_HIP(hipMemcpyAsync(out, t->trie.nodes.p, nn * 16, hipMemcpyDeviceToHost, ctx->stream));
    const uint32_t W = (t->trie.n + 63) / 64;
    if (out_sets && t->trie.m)
      SW_CHECK_HIP(hipMemcpyAsync(out_sets, t->trie.sets.p, t->trie.m * W * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (out_tags && t->trie.m)
      SW_CHECK_HIP(hipMemcpyAsync(out_tags, t->trie.tags.p, t->trie.m * 4, hipMemcpyDeviceToHost, ctx->stream));
    SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return SW_OK;
  } catch (const sw::Error& e) {
    sw::set_error(e.msg);
    return e.code;
  }
}

extern "C" void sw_trie_destroy(sw_trie* t) {
  if (!t) return;
  (void)hipSetDevice(t->ctx->device);
  (void)hipStreamSynchronize(t->ctx->stream);
  delete t;
}
