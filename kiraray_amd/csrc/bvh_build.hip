// bvh_build.hip - binned-SAH BVH builder on the GPU, producing the 4-wide
// layout of bvh.h (the layout the traversal kernels read).
//
// Same algorithm and decisions as the host builder (bvh.cpp; binned SAH
// after Wald, IEEE RT 2007): per node the centroids are binned into 16 bins
// per axis, the split with the lowest cost nodeCost + (A_L N_L + A_R N_R) / A
// is taken (first minimum in axis-major, bin-minor order), a node of at most
// maxLeafSize triangles becomes a leaf when that is cheaper, a node whose
// centroids coincide is split at the middle (or kept as a leaf of <= 8).
// The float arithmetic is the host's, operation for operation, so both build
// the same binary tree except where the host's unstable partition order
// reaches a middle split.
//
// GPU organisation:
//  * large nodes (> kSmall references) are processed level by level, all
//    nodes of a level together: bounds, bins (LDS per block, merged with
//    integer atomics on order-preserving float encodings), split choice, and
//    a stable partition (per-block counts, a per-node scan, a scatter);
//  * every node of at most kSmall references roots a subtree that one wave
//    builds alone: bins in the wave's LDS, ballots for the stable partition,
//    a depth-first stack in global memory;
//  * the binary tree is collapsed into 4-wide nodes level by level (open the
//    child with the largest surface area until four children), nodes of a
//    level numbered by a prefix sum, children of a node contiguous;
//  * leaf triangles are stored in reference order, so a binary leaf
//    [first, first + count) is directly its triangle range.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "bvh.h"
#include "bvh_build.h"

namespace krr {
namespace bvhdev {

constexpr int NB	   = 16;   // SAH bins per axis (the host builder's default)
constexpr int kSmall   = 1024; // nodes of at most this many references are built by one wave
constexpr int kChunk   = 2048; // references per block in the level-synchronous phase
constexpr int kThreads = 256;

// Order-preserving float <-> int map: a < b (floats) <=> ordF(a) < ordF(b).
__host__ __device__ inline int ordF(float f) {
	int i;
	std::memcpy(&i, &f, 4);
	return i >= 0 ? i : i ^ 0x7fffffff;
}
__host__ __device__ inline float unordF(int i) {
	int j = i >= 0 ? i : i ^ 0x7fffffff;
	float f;
	std::memcpy(&f, &j, 4);
	return f;
}

struct Box {
	float lo[3], hi[3];
	__host__ __device__ void reset() {
		for (int a = 0; a < 3; a++) {
			lo[a] = __builtin_inff();
			hi[a] = -__builtin_inff();
		}
	}
	__host__ __device__ void grow(const Box &b) {
		for (int a = 0; a < 3; a++) {
			lo[a] = fminf(lo[a], b.lo[a]);
			hi[a] = fmaxf(hi[a], b.hi[a]);
		}
	}
	// the host builder's Box::area, operation for operation
	__host__ __device__ float area() const {
		float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
		if (!(e0 >= 0) || !(e1 >= 0) || !(e2 >= 0)) return 0.f;
		return 2.f * (e0 * e1 + e1 * e2 + e2 * e0);
	}
};

// Bins of one node: per axis and bin the reference count and the union of
// the primitive boxes, as order-preserving ints.
struct Bins {
	int cnt[3][NB];
	int lo[3][NB][3];
	int hi[3][NB][3];
};

__device__ inline void binsReset(Bins &b, int t, int nt) {
	int *p = &b.cnt[0][0];
	for (int i = t; i < (int) (sizeof(Bins) / 4); i += nt) {
		const int k = i - 3 * NB; // lo block, then hi block
		p[i]		= i < 3 * NB ? 0 : (k < 3 * NB * 3 ? ordF(__builtin_inff()) : ordF(-__builtin_inff()));
	}
}

__device__ inline int binOf(float c, float lo, float scale) { return min(NB - 1, (int) ((c - lo) * scale)); }

struct Split {
	int axis, bin; // axis < 0: no binned split
	float cost;
};

// SAH over the bins exactly as the host builder (bvh.cpp Builder2::build):
// suffix boxes from the top bin down, prefix boxes from bin 0 up, strict '<'.
__device__ Split chooseSplit(const Bins &b, const Box &cb, float pA, float nodeCost) {
	Split s{-1, -1, __builtin_inff()};
	for (int a = 0; a < 3; a++) {
		const float ext = cb.hi[a] - cb.lo[a];
		if (!(ext > 0)) continue;
		float rightArea[NB];
		int rightCnt[NB];
		Box acc;
		acc.reset();
		int c = 0;
		for (int k = NB - 1; k >= 1; k--) {
			Box bk;
			for (int e = 0; e < 3; e++) {
				bk.lo[e] = unordF(b.lo[a][k][e]);
				bk.hi[e] = unordF(b.hi[a][k][e]);
			}
			acc.grow(bk);
			c += b.cnt[a][k];
			rightArea[k] = acc.area();
			rightCnt[k]	 = c;
		}
		acc.reset();
		c = 0;
		for (int k = 0; k < NB - 1; k++) {
			Box bk;
			for (int e = 0; e < 3; e++) {
				bk.lo[e] = unordF(b.lo[a][k][e]);
				bk.hi[e] = unordF(b.hi[a][k][e]);
			}
			acc.grow(bk);
			c += b.cnt[a][k];
			if (c == 0 || rightCnt[k + 1] == 0) continue;
			const float cost = nodeCost + (acc.area() * c + rightArea[k + 1] * rightCnt[k + 1]) / fmaxf(pA, 1e-30f);
			if (cost < s.cost) {
				s.cost = cost;
				s.axis = a;
				s.bin  = k;
			}
		}
	}
	return s;
}

// Decision for a node of `count` references: 0 split (binned), 1 leaf,
// 2 split at the middle (bvh.cpp, after the SAH loop).
__device__ inline int decide(const Split &s, int count, int maxLeaf) {
	if (count <= maxLeaf && (s.axis < 0 || (float) count <= s.cost)) return 1;
	if (s.axis >= 0) return 0;
	return count <= 8 ? 1 : 2;
}

struct Prims {
	const float4 *lo, *hi; // primitive boxes (w unused)
};

__device__ inline void primBox(const Prims &P, int p, Box &b, float c[3]) {
	const float4 l = P.lo[p], h = P.hi[p];
	b.lo[0] = l.x, b.lo[1] = l.y, b.lo[2] = l.z;
	b.hi[0] = h.x, b.hi[1] = h.y, b.hi[2] = h.z;
	for (int a = 0; a < 3; a++) c[a] = 0.5f * (b.lo[a] + b.hi[a]);
}

// Binary node: lo.w = left child, hi.w = right child (as int bits), range =
// {first, count}; count > 0 marks a leaf.
struct Nodes2 {
	float4 *lo, *hi;
	int2 *range;
};

__device__ inline void writeNode(const Nodes2 &N, int i, const Box &b, int left, int right, int first, int count) {
	N.lo[i]	   = make_float4(b.lo[0], b.lo[1], b.lo[2], __int_as_float(left));
	N.hi[i]	   = make_float4(b.hi[0], b.hi[1], b.hi[2], __int_as_float(right));
	N.range[i] = make_int2(first, count);
}

// ------------------------------------------------------------ primitives

__global__ void k_prim_boxes(const float *pos, const int *idx, int n, float4 *lo, float4 *hi, int *refs) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	float l[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()}, h[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
	for (int v = 0; v < 3; v++) {
		const int vi = idx[3 * (size_t) i + v];
		for (int a = 0; a < 3; a++) {
			const float x = pos[3 * (size_t) vi + a];
			l[a]		  = fminf(l[a], x);
			h[a]		  = fmaxf(h[a], x);
		}
	}
	lo[i]	= make_float4(l[0], l[1], l[2], 0.f);
	hi[i]	= make_float4(h[0], h[1], h[2], 0.f);
	refs[i] = i;
}

// ------------------------------------------------- large nodes, per level

// Per node of the level: first, count, binary node index; bounds and bins
// accumulated by the blocks of the node.
struct Level {
	const int *first, *count, *node;
	int *box;  // 6 ints per node: lo xyz, hi xyz (ordered)
	int *cbox; // centroid bounds, same layout
	Bins *bins;
	int *decision, *axis, *bin; // split per node
	int *nLeft;					// left references per node
	const int *blockNode;		// node of each block
	const int *blockStart;		// first reference of each block
	int *blockLeft, *blockOff;	// per-block left counts and their exclusive prefix within the node
	const int *blockBegin;		// first block of each node (n + 1 entries)
};

__global__ void k_level_reset(Level L, int nNodes) {
	const int t = blockIdx.x;
	if (t >= nNodes) return;
	for (int e = threadIdx.x; e < 12; e += blockDim.x) {
		const int v = (e % 6) < 3 ? ordF(__builtin_inff()) : ordF(-__builtin_inff());
		if (e < 6) L.box[6 * t + e] = v;
		else L.cbox[6 * t + e - 6] = v;
	}
	binsReset(L.bins[t], threadIdx.x, blockDim.x);
}

__device__ inline float waveMin(float v) {
	for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
	return v;
}
__device__ inline float waveMax(float v) {
	for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
	return v;
}

__global__ void __launch_bounds__(kThreads) k_level_bounds(Level L, Prims P, const int *refs) {
	const int b = blockIdx.x, t = L.blockNode[b];
	const int end = min(L.blockStart[b] + kChunk, L.first[t] + L.count[t]);
	Box bb, cb;
	bb.reset();
	cb.reset();
	for (int i = L.blockStart[b] + (int) threadIdx.x; i < end; i += kThreads) {
		Box pb;
		float c[3];
		primBox(P, refs[i], pb, c);
		bb.grow(pb);
		for (int a = 0; a < 3; a++) {
			cb.lo[a] = fminf(cb.lo[a], c[a]);
			cb.hi[a] = fmaxf(cb.hi[a], c[a]);
		}
	}
	for (int a = 0; a < 3; a++) {
		const float v[4] = {waveMin(bb.lo[a]), waveMax(bb.hi[a]), waveMin(cb.lo[a]), waveMax(cb.hi[a])};
		if (__lane_id() == 0) {
			atomicMin(&L.box[6 * t + a], ordF(v[0]));
			atomicMax(&L.box[6 * t + 3 + a], ordF(v[1]));
			atomicMin(&L.cbox[6 * t + a], ordF(v[2]));
			atomicMax(&L.cbox[6 * t + 3 + a], ordF(v[3]));
		}
	}
}

__device__ inline Box loadBox(const int *p) {
	Box b;
	for (int a = 0; a < 3; a++) {
		b.lo[a] = unordF(p[a]);
		b.hi[a] = unordF(p[3 + a]);
	}
	return b;
}

__global__ void __launch_bounds__(kThreads) k_level_bins(Level L, Prims P, const int *refs) {
	__shared__ Bins sb;
	const int b = blockIdx.x, t = L.blockNode[b];
	binsReset(sb, threadIdx.x, kThreads);
	__syncthreads();
	const Box cb  = loadBox(L.cbox + 6 * t);
	const int end = min(L.blockStart[b] + kChunk, L.first[t] + L.count[t]);
	float scale[3];
	for (int a = 0; a < 3; a++) {
		const float ext = cb.hi[a] - cb.lo[a];
		scale[a]		= ext > 0 ? NB / ext : 0.f;
	}
	for (int i = L.blockStart[b] + (int) threadIdx.x; i < end; i += kThreads) {
		Box pb;
		float c[3];
		primBox(P, refs[i], pb, c);
		for (int a = 0; a < 3; a++) {
			if (!(scale[a] > 0)) continue;
			const int k = binOf(c[a], cb.lo[a], scale[a]);
			atomicAdd(&sb.cnt[a][k], 1);
			for (int e = 0; e < 3; e++) {
				atomicMin(&sb.lo[a][k][e], ordF(pb.lo[e]));
				atomicMax(&sb.hi[a][k][e], ordF(pb.hi[e]));
			}
		}
	}
	__syncthreads();
	Bins &g = L.bins[t];
	for (int i = threadIdx.x; i < 3 * NB; i += kThreads) {
		const int a = i / NB, k = i % NB;
		if (sb.cnt[a][k] == 0) continue;
		atomicAdd(&g.cnt[a][k], sb.cnt[a][k]);
		for (int e = 0; e < 3; e++) {
			atomicMin(&g.lo[a][k][e], sb.lo[a][k][e]);
			atomicMax(&g.hi[a][k][e], sb.hi[a][k][e]);
		}
	}
}

__global__ void k_level_split(Level L, int nNodes, float nodeCost, int maxLeaf) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nNodes) return;
	const Box bb = loadBox(L.box + 6 * t), cb = loadBox(L.cbox + 6 * t);
	const Split s = chooseSplit(L.bins[t], cb, bb.area(), nodeCost);
	const int d	  = decide(s, L.count[t], maxLeaf);
	L.decision[t] = d;
	L.axis[t]	  = s.axis;
	L.bin[t]	  = s.bin;
}

__device__ inline bool goesLeft(const Level &L, int t, const Box &cb, float scale, int pos, float c) {
	if (L.decision[t] != 0) return pos < L.first[t] + L.count[t] / 2;
	return binOf(c, cb.lo[L.axis[t]], scale) <= L.bin[t];
}

__device__ inline float splitScale(const Level &L, int t, const Box &cb) {
	const int a = L.axis[t] < 0 ? 0 : L.axis[t];
	const float ext = cb.hi[a] - cb.lo[a];
	return ext > 0 ? NB / ext : 0.f;
}

__global__ void __launch_bounds__(kThreads) k_level_count(Level L, Prims P, const int *refs) {
	__shared__ int part[kThreads / 64];
	const int b = blockIdx.x, t = L.blockNode[b];
	const Box cb	  = loadBox(L.cbox + 6 * t);
	const float scale = splitScale(L, t, cb);
	const int ax	  = L.axis[t] < 0 ? 0 : L.axis[t];
	const int end	  = min(L.blockStart[b] + kChunk, L.first[t] + L.count[t]);
	int n = 0;
	for (int i = L.blockStart[b] + (int) threadIdx.x; i < end; i += kThreads) {
		Box pb;
		float c[3];
		primBox(P, refs[i], pb, c);
		n += goesLeft(L, t, cb, scale, i, c[ax]) ? 1 : 0;
	}
	for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
	if (__lane_id() == 0) part[threadIdx.x / 64] = n;
	__syncthreads();
	if (threadIdx.x == 0) {
		int s = 0;
		for (int w = 0; w < kThreads / 64; w++) s += part[w];
		L.blockLeft[b] = s;
	}
}

// Exclusive prefix of the block counts of each node (blocks of a node are
// consecutive), and the node's left total.
__global__ void k_level_scan(Level L, int nNodes) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nNodes) return;
	int s = 0;
	for (int b = L.blockBegin[t]; b < L.blockBegin[t + 1]; b++) {
		L.blockOff[b] = s;
		s += L.blockLeft[b];
	}
	L.nLeft[t] = s;
}

// Stable partition: thread j of the block owns references [8j, 8j + 8) of
// the block's chunk, in order.
__global__ void __launch_bounds__(kThreads) k_level_scatter(Level L, Prims P, const int *refs, int *tmp) {
	constexpr int per = kChunk / kThreads;
	__shared__ int lefts[kThreads];
	const int b = blockIdx.x, t = L.blockNode[b];
	const Box cb	  = loadBox(L.cbox + 6 * t);
	const float scale = splitScale(L, t, cb);
	const int ax	  = L.axis[t] < 0 ? 0 : L.axis[t];
	const int first = L.first[t], end = min(L.blockStart[b] + kChunk, first + L.count[t]);
	const int i0 = L.blockStart[b] + (int) threadIdx.x * per;
	int r[per];
	bool left[per];
	int nl = 0;
	for (int k = 0; k < per; k++) {
		const int i = i0 + k;
		r[k]		= i < end ? refs[i] : -1;
		left[k]		= false;
		if (i < end) {
			Box pb;
			float c[3];
			primBox(P, r[k], pb, c);
			left[k] = goesLeft(L, t, cb, scale, i, c[ax]);
			nl += left[k] ? 1 : 0;
		}
	}
	lefts[threadIdx.x] = nl;
	__syncthreads();
	// exclusive prefix over the threads (Hillis-Steele in LDS)
	for (int o = 1; o < kThreads; o <<= 1) {
		const int v = threadIdx.x >= (unsigned) o ? lefts[threadIdx.x - o] : 0;
		__syncthreads();
		lefts[threadIdx.x] += v;
		__syncthreads();
	}
	const int leftBefore  = lefts[threadIdx.x] - nl; // lefts of earlier threads of this block
	const int blockBefore = i0 - L.blockStart[b];	  // references of earlier threads of this block
	int lp = first + L.blockOff[b] + leftBefore;
	int rp = first + L.nLeft[t] + (L.blockStart[b] - first - L.blockOff[b]) + (blockBefore - leftBefore);
	for (int k = 0; k < per; k++) {
		if (i0 + k >= end) break;
		if (left[k]) tmp[lp++] = r[k];
		else tmp[rp++] = r[k];
	}
}

__global__ void __launch_bounds__(kThreads) k_level_copy(Level L, const int *tmp, int *refs) {
	const int b = blockIdx.x, t = L.blockNode[b];
	const int end = min(L.blockStart[b] + kChunk, L.first[t] + L.count[t]);
	for (int i = L.blockStart[b] + (int) threadIdx.x; i < end; i += kThreads) refs[i] = tmp[i];
}

// Binary nodes of the level: box, and children (childBase[t], +1) or a leaf.
__global__ void k_level_nodes(Level L, int nNodes, const int *childBase, Nodes2 N) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nNodes) return;
	const Box bb = loadBox(L.box + 6 * t);
	if (L.decision[t] == 1) writeNode(N, L.node[t], bb, -1, -1, L.first[t], L.count[t]);
	else writeNode(N, L.node[t], bb, childBase[t], childBase[t] + 1, 0, 0);
}

// --------------------------------------------- small subtrees, one per wave

struct SmallTask {
	int node, first, count, nodeBase; // descendants take nodeBase, nodeBase + 1, ...
};

// One wave (one 64-thread block) builds the subtree of one task depth first,
// with the task's references, the partition scratch and the stack of pending
// right children in LDS.  Per node: one pass bins the references, one pass
// partitions them (the left count comes from the bins) and accumulates both
// children's bounds; the left child is processed next with those bounds, a
// popped right child first recomputes its own.
__device__ inline void waveBox(Box &b) {
	for (int a = 0; a < 3; a++) {
		b.lo[a] = waveMin(b.lo[a]);
		b.hi[a] = waveMax(b.hi[a]);
	}
}

__global__ void __launch_bounds__(64) k_small(const SmallTask *tasks, int nTasks, Prims P, int *refs, Nodes2 N, float nodeCost, int maxLeaf) {
	__shared__ int sRef[kSmall], sTmp[kSmall];
	__shared__ int2 sStack[kSmall]; // {node, first << 11 | count}, first relative to the task
	__shared__ Bins B;
	const int ti = blockIdx.x, lane = threadIdx.x;
	if (ti >= nTasks) return;
	const SmallTask tk = tasks[ti];
	for (int i = lane; i < tk.count; i += 64) sRef[i] = refs[tk.first + i];
	__syncthreads();
	int sp = 0, nextNode = tk.nodeBase;
	int node = tk.node, first = 0, count = tk.count; // first relative to the task
	bool haveBounds = false;
	Box bb, cb;
	while (true) {
		if (!haveBounds) { // bounds and centroid bounds of the node
			bb.reset();
			cb.reset();
			for (int i = first + lane; i < first + count; i += 64) {
				Box pb;
				float c[3];
				primBox(P, sRef[i], pb, c);
				bb.grow(pb);
				for (int a = 0; a < 3; a++) {
					cb.lo[a] = fminf(cb.lo[a], c[a]);
					cb.hi[a] = fmaxf(cb.hi[a], c[a]);
				}
			}
			waveBox(bb);
			waveBox(cb);
		}
		int d = 1;
		Split s{-1, -1, 0.f};
		float scale[3];
		if (count > 1) {
			binsReset(B, lane, 64);
			__syncthreads();
			for (int a = 0; a < 3; a++) {
				const float ext = cb.hi[a] - cb.lo[a];
				scale[a]		= ext > 0 ? NB / ext : 0.f;
			}
			for (int i = first + lane; i < first + count; i += 64) {
				Box pb;
				float c[3];
				primBox(P, sRef[i], pb, c);
				for (int a = 0; a < 3; a++) {
					if (!(scale[a] > 0)) continue;
					const int k = binOf(c[a], cb.lo[a], scale[a]);
					atomicAdd(&B.cnt[a][k], 1);
					for (int e = 0; e < 3; e++) {
						atomicMin(&B.lo[a][k][e], ordF(pb.lo[e]));
						atomicMax(&B.hi[a][k][e], ordF(pb.hi[e]));
					}
				}
			}
			__syncthreads();
			s = chooseSplit(B, cb, bb.area(), nodeCost); // every lane, same result
			d = decide(s, count, maxLeaf);
		}
		if (d == 1) {
			if (lane == 0) writeNode(N, node, bb, -1, -1, tk.first + first, count);
			__syncthreads(); // bins read before the next node resets them
			if (sp == 0) break;
			const int2 e = sStack[--sp];
			node		 = e.x;
			first		 = e.y >> 11;
			count		 = e.y & 2047;
			haveBounds	 = false;
			continue;
		}
		// left count from the bins (binned split) or half (middle split)
		int nLeft = count / 2;
		if (d == 0) {
			nLeft = 0;
			for (int k = 0; k <= s.bin; k++) nLeft += B.cnt[s.axis][k];
		}
		__syncthreads(); // bins read before the next node resets them
		// stable partition into sTmp, with the left child's bounds
		Box lb, lc;
		lb.reset();
		lc.reset();
		int nl = 0, nr = 0;
		for (int base = first; base < first + count; base += 64) {
			const int i	 = base + lane;
			const bool v = i < first + count;
			bool left	 = false;
			int r		 = 0;
			Box pb;
			float c[3];
			if (v) {
				r = sRef[i];
				primBox(P, r, pb, c);
				left = d == 2 ? i < first + count / 2 : binOf(c[s.axis], cb.lo[s.axis], scale[s.axis]) <= s.bin;
				if (left) {
					lb.grow(pb);
					for (int a = 0; a < 3; a++) {
						lc.lo[a] = fminf(lc.lo[a], c[a]);
						lc.hi[a] = fmaxf(lc.hi[a], c[a]);
					}
				}
			}
			const uint64_t lm = __ballot(v && left), rm = __ballot(v && !left);
			const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
			const int slot = v ? (left ? first + nl + __popcll(lm & below) : first + nLeft + nr + __popcll(rm & below)) : -1;
			if (slot >= first && slot < first + count) sTmp[slot] = r; // always in range when nl ends equal to nLeft
			nl += __popcll(lm);
			nr += __popcll(rm);
		}
		__syncthreads();
		for (int i = first + lane; i < first + count; i += 64) sRef[i] = sTmp[i];
		__syncthreads();
		// nl == nLeft by construction; a side is never empty (binned splits
		// skip empty sides, middle splits halve at least 9 references)
		const int l = nextNode, rgt = nextNode + 1;
		nextNode += 2;
		if (lane == 0) {
			writeNode(N, node, bb, l, rgt, 0, 0);
			sStack[sp] = make_int2(rgt, ((first + nLeft) << 11) | (count - nLeft));
		}
		sp++;
		__syncthreads();
		waveBox(lb);
		waveBox(lc);
		node	   = l;
		count	   = nLeft;
		bb		   = lb;
		cb		   = lc;
		haveBounds = true;
	}
	__syncthreads();
	for (int i = lane; i < tk.count; i += 64) refs[tk.first + i] = sRef[i];
}

// ------------------------------------------------------------- collapse

__device__ inline void padBox(float &lo, float &hi) {
	const float pad = (fabsf(lo) + fabsf(hi) + (hi - lo)) * 9.5367431640625e-07f; // 2^-20, as the host builder
	lo -= pad;
	hi += pad;
}

__device__ inline float nodeArea(const Nodes2 &N, int i) {
	const float4 l = N.lo[i], h = N.hi[i];
	Box b;
	b.lo[0] = l.x, b.lo[1] = l.y, b.lo[2] = l.z;
	b.hi[0] = h.x, b.hi[1] = h.y, b.hi[2] = h.z;
	return b.area();
}

// Children of the 4-wide node made from binary node b: open the inner child
// with the largest area (first maximum) until four children or only leaves.
__device__ int collapseKids(const Nodes2 &N, int b, int kids[4]) {
	int nk = 0;
	if (N.range[b].y > 0) {
		kids[nk++] = b;
		return nk;
	}
	kids[nk++] = __float_as_int(N.lo[b].w);
	kids[nk++] = __float_as_int(N.hi[b].w);
	while (nk < 4) {
		int best = -1;
		float ba = -1.f;
		for (int k = 0; k < nk; k++) {
			if (N.range[kids[k]].y > 0) continue;
			const float a = nodeArea(N, kids[k]);
			if (a > ba) {
				ba	 = a;
				best = k;
			}
		}
		if (best < 0) break;
		const int c = kids[best];
		kids[best]	= __float_as_int(N.lo[c].w);
		kids[nk++]	= __float_as_int(N.hi[c].w);
	}
	return nk;
}

__global__ void k_collapse_count(Nodes2 N, const int *level, int n, int *inner) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int kids[4];
	const int nk = collapseKids(N, level[i], kids);
	int c		 = 0;
	for (int k = 0; k < nk; k++) c += N.range[kids[k]].y > 0 ? 0 : 1;
	inner[i] = c;
}

// Exclusive prefix sum of `n` ints in one block of 1024 threads (chunked).
__global__ void __launch_bounds__(1024) k_scan_1block(const int *in, int *out, int n, int *total) {
	__shared__ int part[1024];
	const int t		= threadIdx.x;
	const int per	= (n + 1023) / 1024;
	const int begin = min(n, t * per), end = min(n, begin + per);
	int s = 0;
	for (int i = begin; i < end; i++) s += in[i];
	part[t] = s;
	__syncthreads();
	for (int o = 1; o < 1024; o <<= 1) {
		const int v = t >= o ? part[t - o] : 0;
		__syncthreads();
		part[t] += v;
		__syncthreads();
	}
	int run = part[t] - s;
	for (int i = begin; i < end; i++) {
		const int v = in[i];
		out[i]		= run;
		run += v;
	}
	if (t == 1023) *total = part[1023];
}

__global__ void k_collapse_emit(Nodes2 N, const int *level, int n, int levelBase, const int *off, int nextBase, int *nextLevel,
								Bvh4Node *out) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int kids[4];
	const int nk = collapseKids(N, level[i], kids);
	Bvh4Node node;
	for (int c = 0; c < 4; c++) {
		node.lox[c] = node.loy[c] = node.loz[c] = __builtin_inff();
		node.hix[c] = node.hiy[c] = node.hiz[c] = -__builtin_inff();
		node.child[c] = -1;
		node.pad[c]	  = 0;
	}
	int j = 0;
	for (int c = 0; c < nk; c++) {
		const int k	   = kids[c];
		const float4 l = N.lo[k], h = N.hi[k];
		float lx = l.x, hx = h.x, ly = l.y, hy = h.y, lz = l.z, hz = h.z;
		padBox(lx, hx);
		padBox(ly, hy);
		padBox(lz, hz);
		node.lox[c] = lx;
		node.hix[c] = hx;
		node.loy[c] = ly;
		node.hiy[c] = hy;
		node.loz[c] = lz;
		node.hiz[c] = hz;
		const int2 r = N.range[k];
		if (r.y > 0) node.child[c] = ~(r.x * 8 + (r.y - 1));
		else {
			const int slot	= off[i] + j++;
			node.child[c]	= nextBase + slot;
			nextLevel[slot] = k;
		}
	}
	out[levelBase + i] = node;
}

__global__ void k_leaf_tris(const float *pos, const int *idx, const unsigned char *nullFlag, const int *refs, int n, LeafTri *tris) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int p = refs[i];
	LeafTri t;
	for (int a = 0; a < 3; a++) {
		t.v0[a] = pos[3 * (size_t) idx[3 * (size_t) p] + a];
		t.v1[a] = pos[3 * (size_t) idx[3 * (size_t) p + 1] + a];
		t.v2[a] = pos[3 * (size_t) idx[3 * (size_t) p + 2] + a];
	}
	t.prim	= p;
	t.flags = nullFlag ? (int32_t) (nullFlag[p] & (kLeafNull | kLeafAlpha)) : 0;
	t.pad	= 0;
	tris[i] = t;
}

// ------------------------------------------------------------------ host

struct Buf {
	void *p = nullptr;
	~Buf() {
		if (p) (void) hipFree(p);
	}
	hipError_t alloc(size_t bytes) {
		if (p) (void) hipFree(p);
		p = nullptr;
		return hipMalloc(&p, std::max<size_t>(bytes, 16));
	}
	template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

#define BVHD_HIP(x)                                                                                                                        \
	do {                                                                                                                                   \
		hipError_t e_ = (x);                                                                                                               \
		if (e_ != hipSuccess) {                                                                                                            \
			err = std::string(#x) + ": " + hipGetErrorString(e_);                                                                          \
			return -1;                                                                                                                     \
		}                                                                                                                                  \
	} while (0)

template <class T> hipError_t upload(Buf &b, const std::vector<T> &v, hipStream_t st) {
	hipError_t e = b.alloc(v.size() * sizeof(T));
	if (e != hipSuccess || v.empty()) return e;
	e = hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
	if (e != hipSuccess) return e;
	return hipStreamSynchronize(st); // v may go out of scope after the call
}

} // namespace bvhdev

using namespace bvhdev;

int buildBvh4Device(const float *pos, int64_t nVerts, const int32_t *idx, int64_t nTris, const uint8_t *nullFlag, const BvhBuildOptions &opt,
					DeviceBvh &out, std::string &err) {
	const auto t0 = std::chrono::steady_clock::now();
	out			  = DeviceBvh{};
	if (nTris <= 0 || nTris > (int64_t) 0x0fffffff) {
		err = "device BVH build needs 1 .. 2^28 triangles";
		return -1;
	}
	const int n		  = (int) nTris;
	const int maxLeaf = std::max(1, std::min(8, opt.maxLeafSize));
	const float cost  = std::max(0.f, opt.nodeCost);
	hipStream_t st	  = nullptr;
	BVHD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
	struct StreamGuard {
		hipStream_t s;
		~StreamGuard() { (void) hipStreamDestroy(s); }
	} guard{st};
	Buf dPos, dIdx, dNull, pLo, pHi, refs, tmp, nLo, nHi, nRange;
	BVHD_HIP(dPos.alloc((size_t) nVerts * 3 * sizeof(float)));
	BVHD_HIP(dIdx.alloc((size_t) n * 3 * sizeof(int)));
	BVHD_HIP(dNull.alloc((size_t) n));
	BVHD_HIP(hipMemcpyAsync(dPos.p, pos, (size_t) nVerts * 3 * sizeof(float), hipMemcpyHostToDevice, st));
	BVHD_HIP(hipMemcpyAsync(dIdx.p, idx, (size_t) n * 3 * sizeof(int), hipMemcpyHostToDevice, st));
	if (nullFlag) BVHD_HIP(hipMemcpyAsync(dNull.p, nullFlag, (size_t) n, hipMemcpyHostToDevice, st));
	else BVHD_HIP(hipMemsetAsync(dNull.p, 0, (size_t) n, st));
	BVHD_HIP(hipStreamSynchronize(st));
	const auto tUp = std::chrono::steady_clock::now();
	BVHD_HIP(pLo.alloc((size_t) n * sizeof(float4)));
	BVHD_HIP(pHi.alloc((size_t) n * sizeof(float4)));
	BVHD_HIP(refs.alloc((size_t) n * sizeof(int)));
	BVHD_HIP(tmp.alloc((size_t) n * sizeof(int)));
	const size_t maxNodes2 = 2 * (size_t) n;
	BVHD_HIP(nLo.alloc(maxNodes2 * sizeof(float4)));
	BVHD_HIP(nHi.alloc(maxNodes2 * sizeof(float4)));
	BVHD_HIP(nRange.alloc(maxNodes2 * sizeof(int2)));
	const Prims P{pLo.as<float4>(), pHi.as<float4>()};
	const Nodes2 N{nLo.as<float4>(), nHi.as<float4>(), nRange.as<int2>()};
	hipLaunchKernelGGL(k_prim_boxes, dim3((n + 255) / 256), dim3(256), 0, st, dPos.as<float>(), dIdx.as<int>(), n, pLo.as<float4>(),
					   pHi.as<float4>(), refs.as<int>());
	BVHD_HIP(hipGetLastError());

	// ---- large nodes, level by level (host keeps the node list)
	std::vector<int> lFirst{0}, lCount{n}, lNode{0};
	std::vector<SmallTask> small;
	int nextNode = 1;
	Buf bFirst, bCount, bNode, bBox, bCbox, bBins, bDec, bAxis, bBin, bLeft, bBlockNode, bBlockStart, bBlockLeft, bBlockOff, bBlockBegin,
		bChild;
	int levels = 0;
	while (!lCount.empty()) {
		// nodes of this level that are small go to the per-wave phase
		std::vector<int> f, c, nd;
		for (size_t t = 0; t < lCount.size(); t++) {
			if (lCount[t] <= kSmall) small.push_back(SmallTask{lNode[t], lFirst[t], lCount[t], 0});
			else {
				f.push_back(lFirst[t]);
				c.push_back(lCount[t]);
				nd.push_back(lNode[t]);
			}
		}
		if (f.empty()) break;
		levels++;
		const int T = (int) f.size();
		std::vector<int> blockNode, blockStart, blockBegin(T + 1);
		for (int t = 0; t < T; t++) {
			blockBegin[t] = (int) blockNode.size();
			for (int s = f[t]; s < f[t] + c[t]; s += kChunk) {
				blockNode.push_back(t);
				blockStart.push_back(s);
			}
		}
		blockBegin[T] = (int) blockNode.size();
		const int B	  = (int) blockNode.size();
		BVHD_HIP(upload(bFirst, f, st));
		BVHD_HIP(upload(bCount, c, st));
		BVHD_HIP(upload(bNode, nd, st));
		BVHD_HIP(upload(bBlockNode, blockNode, st));
		BVHD_HIP(upload(bBlockStart, blockStart, st));
		BVHD_HIP(upload(bBlockBegin, blockBegin, st));
		BVHD_HIP(bBox.alloc((size_t) T * 6 * sizeof(int)));
		BVHD_HIP(bCbox.alloc((size_t) T * 6 * sizeof(int)));
		BVHD_HIP(bBins.alloc((size_t) T * sizeof(Bins)));
		BVHD_HIP(bDec.alloc((size_t) T * sizeof(int)));
		BVHD_HIP(bAxis.alloc((size_t) T * sizeof(int)));
		BVHD_HIP(bBin.alloc((size_t) T * sizeof(int)));
		BVHD_HIP(bLeft.alloc((size_t) T * sizeof(int)));
		BVHD_HIP(bBlockLeft.alloc((size_t) B * sizeof(int)));
		BVHD_HIP(bBlockOff.alloc((size_t) B * sizeof(int)));
		Level L{bFirst.as<int>(),	 bCount.as<int>(),	  bNode.as<int>(),		bBox.as<int>(),		   bCbox.as<int>(),
				bBins.as<Bins>(),	 bDec.as<int>(),	  bAxis.as<int>(),		bBin.as<int>(),		   bLeft.as<int>(),
				bBlockNode.as<int>(), bBlockStart.as<int>(), bBlockLeft.as<int>(), bBlockOff.as<int>(), bBlockBegin.as<int>()};
		hipLaunchKernelGGL(k_level_reset, dim3(T), dim3(256), 0, st, L, T);
		hipLaunchKernelGGL(k_level_bounds, dim3(B), dim3(kThreads), 0, st, L, P, (const int *) refs.as<int>());
		hipLaunchKernelGGL(k_level_bins, dim3(B), dim3(kThreads), 0, st, L, P, (const int *) refs.as<int>());
		hipLaunchKernelGGL(k_level_split, dim3((T + 63) / 64), dim3(64), 0, st, L, T, cost, maxLeaf);
		hipLaunchKernelGGL(k_level_count, dim3(B), dim3(kThreads), 0, st, L, P, (const int *) refs.as<int>());
		hipLaunchKernelGGL(k_level_scan, dim3((T + 63) / 64), dim3(64), 0, st, L, T);
		hipLaunchKernelGGL(k_level_scatter, dim3(B), dim3(kThreads), 0, st, L, P, (const int *) refs.as<int>(), tmp.as<int>());
		hipLaunchKernelGGL(k_level_copy, dim3(B), dim3(kThreads), 0, st, L, (const int *) tmp.as<int>(), refs.as<int>());
		BVHD_HIP(hipGetLastError());
		std::vector<int> dec(T), left(T);
		BVHD_HIP(hipMemcpyAsync(dec.data(), bDec.p, T * sizeof(int), hipMemcpyDeviceToHost, st));
		BVHD_HIP(hipMemcpyAsync(left.data(), bLeft.p, T * sizeof(int), hipMemcpyDeviceToHost, st));
		BVHD_HIP(hipStreamSynchronize(st));
		std::vector<int> childBase(T, -1), nf, nc, nn;
		for (int t = 0; t < T; t++) {
			if (dec[t] == 1) continue; // not reached for nodes above kSmall (> 8 references)
			int nl = dec[t] == 2 ? c[t] / 2 : left[t];
			if (nl <= 0 || nl >= c[t]) {
				err = "device BVH build: empty side in a large-node split";
				return -1;
			}
			childBase[t] = nextNode;
			nextNode += 2;
			nf.push_back(f[t]);
			nc.push_back(nl);
			nn.push_back(childBase[t]);
			nf.push_back(f[t] + nl);
			nc.push_back(c[t] - nl);
			nn.push_back(childBase[t] + 1);
		}
		BVHD_HIP(upload(bChild, childBase, st));
		hipLaunchKernelGGL(k_level_nodes, dim3((T + 63) / 64), dim3(64), 0, st, L, T, (const int *) bChild.as<int>(), N);
		BVHD_HIP(hipGetLastError());
		BVHD_HIP(hipStreamSynchronize(st)); // the level buffers are reallocated next round
		lFirst.swap(nf);
		lCount.swap(nc);
		lNode.swap(nn);
	}
	const auto tLarge = std::chrono::steady_clock::now();
	// ---- small subtrees, one wave each; descendants numbered per subtree
	for (SmallTask &t : small) {
		t.nodeBase = nextNode;
		nextNode += 2 * t.count - 2;
	}
	if ((size_t) nextNode > maxNodes2) {
		err = "device BVH build: node count overflow";
		return -1;
	}
	if (!small.empty()) {
		Buf bSmall;
		BVHD_HIP(upload(bSmall, small, st));
		hipLaunchKernelGGL(k_small, dim3((unsigned) small.size()), dim3(64), 0, st, (const SmallTask *) bSmall.as<SmallTask>(),
						   (int) small.size(), P, refs.as<int>(), N, cost, maxLeaf);
		BVHD_HIP(hipGetLastError());
		BVHD_HIP(hipStreamSynchronize(st));
	}
	const auto tSmall = std::chrono::steady_clock::now();
	// ---- collapse to 4-wide nodes, level by level
	Buf lvl[2], inner, off, total, nodes4, tris;
	BVHD_HIP(lvl[0].alloc(maxNodes2 * sizeof(int)));
	BVHD_HIP(lvl[1].alloc(maxNodes2 * sizeof(int)));
	BVHD_HIP(inner.alloc(maxNodes2 * sizeof(int)));
	BVHD_HIP(off.alloc(maxNodes2 * sizeof(int)));
	BVHD_HIP(total.alloc(sizeof(int)));
	// at most one 4-wide node per inner binary node, plus the root
	const size_t cap4 = (size_t) nextNode / 2 + 1;
	BVHD_HIP(nodes4.alloc(cap4 * sizeof(Bvh4Node)));
	const int zero = 0;
	BVHD_HIP(hipMemcpyAsync(lvl[0].p, &zero, sizeof(int), hipMemcpyHostToDevice, st));
	int levelN = 1, levelBase = 0, depth = 0, cur = 0;
	while (levelN > 0) {
		depth++;
		// a well-formed tree never trips these; they keep a malformed one from
		// writing past the node array or looping
		if (depth > 64 || (size_t) levelBase + (size_t) levelN > cap4 || (size_t) levelN > maxNodes2) {
			err = "device BVH build: malformed tree in the 4-wide collapse";
			return -1;
		}
		const unsigned g = (unsigned) ((levelN + 255) / 256);
		hipLaunchKernelGGL(k_collapse_count, dim3(g), dim3(256), 0, st, N, (const int *) lvl[cur].as<int>(), levelN, inner.as<int>());
		hipLaunchKernelGGL(k_scan_1block, dim3(1), dim3(1024), 0, st, (const int *) inner.as<int>(), off.as<int>(), levelN, total.as<int>());
		const int nextBase = levelBase + levelN;
		hipLaunchKernelGGL(k_collapse_emit, dim3(g), dim3(256), 0, st, N, (const int *) lvl[cur].as<int>(), levelN, levelBase,
						   (const int *) off.as<int>(), nextBase, lvl[cur ^ 1].as<int>(), nodes4.as<Bvh4Node>());
		BVHD_HIP(hipGetLastError());
		int nextN = 0;
		BVHD_HIP(hipMemcpyAsync(&nextN, total.p, sizeof(int), hipMemcpyDeviceToHost, st));
		BVHD_HIP(hipStreamSynchronize(st));
		levelBase = nextBase;
		levelN	  = nextN;
		cur ^= 1;
	}
	BVHD_HIP(tris.alloc((size_t) n * sizeof(LeafTri)));
	hipLaunchKernelGGL(k_leaf_tris, dim3((n + 255) / 256), dim3(256), 0, st, (const float *) dPos.as<float>(), (const int *) dIdx.as<int>(),
					   (const unsigned char *) dNull.as<unsigned char>(), (const int *) refs.as<int>(), n, tris.as<LeafTri>());
	BVHD_HIP(hipGetLastError());
	BVHD_HIP(hipStreamSynchronize(st));
	out.nodes	   = nodes4.p;
	out.tris	   = tris.p;
	nodes4.p	   = nullptr;
	tris.p		   = nullptr;
	out.nNodes	   = levelBase;
	out.nTris	   = n;
	out.nBinary	   = nextNode;
	out.maxDepth   = depth;
	out.largeLevels = levels;
	out.smallTrees = (int64_t) small.size();
	const auto t1  = std::chrono::steady_clock::now();
	out.seconds	   = std::chrono::duration<double>(t1 - t0).count();
	out.buildSeconds = std::chrono::duration<double>(t1 - tUp).count();
	out.largeSeconds = std::chrono::duration<double>(tLarge - tUp).count();
	out.smallSeconds = std::chrono::duration<double>(tSmall - tLarge).count();
	out.collapseSeconds = std::chrono::duration<double>(t1 - tSmall).count();
	return 0;
}

} // namespace krr
