// bvh.cpp - host binned-SAH BVH builder, collapsed to the 4-wide layout the
// traversal kernels read (bvh.h).
//
// Binned SAH after Wald, "On fast Construction of SAH-based Bounding Volume
// Hierarchies" (IEEE RT 2007): per node, primitive centroids are binned along
// each axis and the split with the lowest SAH cost is taken.  Large subtrees
// are built on separate host threads.  The binary tree is then collapsed to a
// 4-wide tree by repeatedly opening the child with the largest surface area.
#include "bvh.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace krr {

namespace {

struct Box {
	float lo[3], hi[3];
	void reset() {
		for (int a = 0; a < 3; a++) {
			lo[a] = std::numeric_limits<float>::infinity();
			hi[a] = -std::numeric_limits<float>::infinity();
		}
	}
	void grow(const Box &b) {
		for (int a = 0; a < 3; a++) {
			lo[a] = std::min(lo[a], b.lo[a]);
			hi[a] = std::max(hi[a], b.hi[a]);
		}
	}
	float area() const {
		float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
		if (!(e0 >= 0) || !(e1 >= 0) || !(e2 >= 0)) return 0.f;
		return 2.f * (e0 * e1 + e1 * e2 + e2 * e0);
	}
};

struct Node2 {
	Box box;
	int left = -1, right = -1;
	int first = 0, count = 0; // leaf if count > 0
};

struct Builder2 {
	const float *pos;
	const int32_t *idx;
	int64_t n;
	BvhBuildOptions opt;
	std::vector<Box> primBox;
	std::vector<float> cen; // 3 per primitive
	std::vector<int> refs;
	std::vector<Node2> nodes;
	std::atomic<int> nodeCount{0};
	std::atomic<int> threadsLeft{0};

	int alloc2() { return nodeCount.fetch_add(2); }

	void build(int nodeIdx, int first, int count, int depth) {
		Node2 &nd = nodes[nodeIdx];
		Box b, cb;
		b.reset();
		cb.reset();
		for (int i = first; i < first + count; i++) {
			int p = refs[i];
			b.grow(primBox[p]);
			for (int a = 0; a < 3; a++) {
				cb.lo[a] = std::min(cb.lo[a], cen[3 * p + a]);
				cb.hi[a] = std::max(cb.hi[a], cen[3 * p + a]);
			}
		}
		nd.box = b;
		if (count == 1) {
			nd.first = first;
			nd.count = count;
			return;
		}
		const int NB	= std::max(2, std::min(64, opt.bins));
		int bestAxis	= -1, bestBin = -1;
		float bestCost	= std::numeric_limits<float>::infinity();
		const float pA	= b.area();
		for (int a = 0; a < 3; a++) {
			float ext = cb.hi[a] - cb.lo[a];
			if (!(ext > 0)) continue;
			int cnt[64];
			Box bb[64];
			for (int k = 0; k < NB; k++) {
				cnt[k] = 0;
				bb[k].reset();
			}
			const float scale = NB / ext;
			for (int i = first; i < first + count; i++) {
				int p = refs[i];
				int k = std::min(NB - 1, (int) ((cen[3 * p + a] - cb.lo[a]) * scale));
				cnt[k]++;
				bb[k].grow(primBox[p]);
			}
			float rightArea[64];
			int rightCnt[64];
			Box acc;
			acc.reset();
			int c = 0;
			for (int k = NB - 1; k >= 1; k--) {
				acc.grow(bb[k]);
				c += cnt[k];
				rightArea[k] = acc.area();
				rightCnt[k]	 = c;
			}
			acc.reset();
			c = 0;
			for (int k = 0; k < NB - 1; k++) {
				acc.grow(bb[k]);
				c += cnt[k];
				if (c == 0 || rightCnt[k + 1] == 0) continue;
				float cost = opt.nodeCost + (acc.area() * c + rightArea[k + 1] * rightCnt[k + 1]) / std::max(pA, 1e-30f);
				if (cost < bestCost) {
					bestCost = cost;
					bestAxis = a;
					bestBin	 = k;
				}
			}
		}
		if (count <= opt.maxLeafSize && (bestAxis < 0 || (float) count <= bestCost)) {
			nd.first = first;
			nd.count = count;
			return;
		}
		int mid;
		if (bestAxis >= 0) {
			float ext	= cb.hi[bestAxis] - cb.lo[bestAxis];
			float scale = NB / ext;
			int *beg	= refs.data() + first;
			int *m		= std::partition(beg, beg + count, [&](int p) {
				 return std::min(NB - 1, (int) ((cen[3 * p + bestAxis] - cb.lo[bestAxis]) * scale)) <= bestBin;
			 });
			mid = first + (int) (m - beg);
			if (mid == first || mid == first + count) mid = first + count / 2;
		} else {
			if (count <= 8) { // identical centroids: accept a larger leaf
				nd.first = first;
				nd.count = count;
				return;
			}
			mid = first + count / 2;
		}
		int l	 = alloc2();
		nd.left	 = l;
		nd.right = l + 1;
		int nl = mid - first, nr = count - nl;
		if (count > 65536 && threadsLeft.fetch_sub(1) > 0) {
			std::thread t([this, l, first, nl, depth]() { build(l, first, nl, depth + 1); });
			build(l + 1, mid, nr, depth + 1);
			t.join();
		} else {
			build(l, first, nl, depth + 1);
			build(l + 1, mid, nr, depth + 1);
		}
	}
};

void padBox(float &lo, float &hi) {
	float pad = (std::fabs(lo) + std::fabs(hi) + (hi - lo)) * 9.5367431640625e-07f; // 2^-20
	lo -= pad;
	hi += pad;
}

struct Collapser {
	const Builder2 &b2;
	const float *pos;
	const int32_t *idx;
	const uint8_t *nullFlag;
	Bvh4Host &out;

	void setChildBox(Bvh4Node &n, int c, const Box &b) {
		float lx = b.lo[0], hx = b.hi[0], ly = b.lo[1], hy = b.hi[1], lz = b.lo[2], hz = b.hi[2];
		padBox(lx, hx);
		padBox(ly, hy);
		padBox(lz, hz);
		n.lox[c] = lx;
		n.hix[c] = hx;
		n.loy[c] = ly;
		n.hiy[c] = hy;
		n.loz[c] = lz;
		n.hiz[c] = hz;
	}

	int32_t emitLeaf(const Node2 &l) {
		int64_t first = (int64_t) out.tris.size();
		for (int i = l.first; i < l.first + l.count; i++) {
			int p = b2.refs[i];
			LeafTri t;
			std::memset(&t, 0, sizeof(t));
			for (int a = 0; a < 3; a++) {
				t.v0[a] = pos[3 * (size_t) idx[3 * (size_t) p] + a];
				t.v1[a] = pos[3 * (size_t) idx[3 * (size_t) p + 1] + a];
				t.v2[a] = pos[3 * (size_t) idx[3 * (size_t) p + 2] + a];
			}
			t.prim	= p;
			t.flags = nullFlag ? (int32_t) (nullFlag[p] & (kLeafNull | kLeafAlpha)) : 0;
			out.tris.push_back(t);
		}
		out.nLeaves++;
		return ~(int32_t) (first * 8 + (l.count - 1));
	}

	int collapse(int n2, int depth) {
		out.maxDepth = std::max(out.maxDepth, depth);
		int me		 = (int) out.nodes.size();
		out.nodes.emplace_back();
		int kids[4];
		int nk		= 0;
		const Node2 &root = b2.nodes[n2];
		if (root.count > 0) kids[nk++] = n2; // a leaf root becomes the only child
		else {
			kids[nk++] = root.left;
			kids[nk++] = root.right;
			while (nk < 4) {
				int best = -1;
				float ba = -1.f;
				for (int k = 0; k < nk; k++) {
					const Node2 &c = b2.nodes[kids[k]];
					if (c.count > 0) continue;
					float a = c.box.area();
					if (a > ba) {
						ba	 = a;
						best = k;
					}
				}
				if (best < 0) break;
				const Node2 &c = b2.nodes[kids[best]];
				kids[best]	   = c.left;
				kids[nk++]	   = c.right;
			}
		}
		Bvh4Node node;
		for (int c = 0; c < 4; c++) {
			node.lox[c] = node.loy[c] = node.loz[c] = std::numeric_limits<float>::infinity();
			node.hix[c] = node.hiy[c] = node.hiz[c] = -std::numeric_limits<float>::infinity();
			node.child[c] = -1;
			node.pad[c]	  = 0;
		}
		for (int c = 0; c < nk; c++) {
			const Node2 &k = b2.nodes[kids[c]];
			setChildBox(node, c, k.box);
			if (k.count > 0) node.child[c] = emitLeaf(k);
			else node.child[c] = collapse(kids[c], depth + 1);
		}
		out.nodes[me] = node;
		return me;
	}
};

} // namespace

void buildBvh4(const float *pos, const int32_t *idx, int64_t nTris, const uint8_t *nullFlag, const BvhBuildOptions &opt,
			   Bvh4Host &out) {
	auto t0 = std::chrono::steady_clock::now();
	out		= Bvh4Host();
	Builder2 b;
	b.pos = pos;
	b.idx = idx;
	b.n	  = nTris;
	b.opt = opt;
	b.opt.maxLeafSize = std::max(1, std::min(8, opt.maxLeafSize));
	b.opt.nodeCost	  = std::max(0.f, opt.nodeCost);
	unsigned hw		  = std::thread::hardware_concurrency();
	int threads		  = opt.threads > 0 ? opt.threads : (int) std::min(16u, std::max(1u, hw));
	b.threadsLeft	  = threads - 1;
	b.primBox.resize((size_t) nTris);
	b.cen.resize(3 * (size_t) nTris);
	b.refs.resize((size_t) nTris);
	for (int64_t i = 0; i < nTris; i++) {
		Box bx;
		bx.reset();
		for (int v = 0; v < 3; v++)
			for (int a = 0; a < 3; a++) {
				float x	  = pos[3 * (size_t) idx[3 * i + v] + a];
				bx.lo[a] = std::min(bx.lo[a], x);
				bx.hi[a] = std::max(bx.hi[a], x);
			}
		b.primBox[i] = bx;
		for (int a = 0; a < 3; a++) b.cen[3 * i + a] = 0.5f * (bx.lo[a] + bx.hi[a]);
		b.refs[i] = (int) i;
	}
	b.nodes.resize((size_t) std::max<int64_t>(1, 2 * nTris));
	b.nodeCount = 1;
	if (nTris > 0) b.build(0, 0, (int) nTris, 0);
	else {
		b.nodes[0].box.reset();
		b.nodes[0].count = 0;
	}
	Collapser c{b, pos, idx, nullFlag, out};
	if (nTris > 0) {
		out.nodes.reserve((size_t) (b.nodeCount.load() / 2 + 1));
		out.tris.reserve((size_t) nTris);
		c.collapse(0, 1);
	} else {
		Bvh4Node node;
		for (int k = 0; k < 4; k++) {
			node.lox[k] = node.loy[k] = node.loz[k] = std::numeric_limits<float>::infinity();
			node.hix[k] = node.hiy[k] = node.hiz[k] = -std::numeric_limits<float>::infinity();
			node.child[k] = -1;
			node.pad[k]	  = 0;
		}
		out.nodes.push_back(node);
		out.maxDepth = 1;
	}
	out.buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace krr
