// Quad generator: pairs edge-adjacent triangles of an indexed mesh into quads, so that the
// rasteriser's setup handles two triangles per quad (src/quad_generator.cpp:8-201).
//
// Same algorithm as the reference:
//   * triangle neighbours across each edge from a directed-edge map (:19-45);
//   * one candidate quad per shared edge, vertices {t0[i], opposite, t0[i+1], t0[i+2]}, scored
//     by squareness = (4 - sum |cos| of consecutive edge directions) / 4 (:8-17, :56-100);
//   * candidates sharing a triangle conflict (:102-110);
//   * greedy maximum independent set: repeatedly take the candidate of least
//     degree - squareness * weight, drop its conflicts and lower their neighbours' degrees
//     (:118-175);
//   * output in the original triangle order; unpaired triangles become (a, b, c, c) (:177-200).
// Ties of the greedy score are broken by candidate index (the reference's heap order belongs
// to a library that is not part of this build).
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "../../include/lucid_c.h"

namespace {

struct V3f {
	float x, y, z;
};

V3f load(const float *p, uint32_t v) { return {p[3 * (size_t)v], p[3 * (size_t)v + 1], p[3 * (size_t)v + 2]}; }
V3f sub(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
float dot(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3f normalize(V3f a) {
	float l = 1.0f / std::sqrt(dot(a, a));
	return {a.x * l, a.y * l, a.z * l};
}

float squareness(const V3f (&c)[4]) {
	V3f e[4];
	for(int i = 0; i < 4; i++)
		e[i] = normalize(sub(c[(i + 1) & 3], c[i]));
	float s = 0.0f;
	for(int i = 0; i < 4; i++)
		s += std::fabs(dot(e[i], e[(i + 1) & 3]));
	return (4.0f - s) * 0.25f;
}

struct Candidate {
	uint32_t tris[2];
	uint32_t verts[4];
	int32_t conflicts[4] = {-1, -1, -1, -1};
	float square;
	void add_conflict(int32_t q) {
		for(int32_t c : conflicts)
			if(c == q)
				return;
		for(int32_t &c : conflicts)
			if(c == -1) {
				c = q;
				return;
			}
	}
	int degree() const {
		int d = 0;
		for(int32_t c : conflicts)
			d += c != -1;
		return d;
	}
};

// Binary min-heap of candidate ids keyed by (score, id), with positions for decrease/increase.
struct Heap {
	std::vector<float> key;
	std::vector<int32_t> heap, pos;
	explicit Heap(size_t n) : key(n), pos(n, -1) { heap.reserve(n); }
	bool less(int32_t a, int32_t b) const {
		return key[a] < key[b] || (key[a] == key[b] && a < b);
	}
	void swap_at(size_t i, size_t j) {
		std::swap(heap[i], heap[j]);
		pos[heap[i]] = (int32_t)i, pos[heap[j]] = (int32_t)j;
	}
	void up(size_t i) {
		while(i > 0) {
			size_t p = (i - 1) / 2;
			if(!less(heap[i], heap[p]))
				break;
			swap_at(i, p);
			i = p;
		}
	}
	void down(size_t i) {
		for(;;) {
			size_t l = 2 * i + 1, r = l + 1, m = i;
			if(l < heap.size() && less(heap[l], heap[m]))
				m = l;
			if(r < heap.size() && less(heap[r], heap[m]))
				m = r;
			if(m == i)
				break;
			swap_at(i, m);
			i = m;
		}
	}
	void insert(int32_t id, float k) {
		key[id] = k;
		pos[id] = (int32_t)heap.size();
		heap.push_back(id);
		up(heap.size() - 1);
	}
	void update(int32_t id, float k) {
		if(pos[id] < 0)
			return;
		key[id] = k;
		up((size_t)pos[id]);
		down((size_t)pos[id]);
	}
	int32_t pop() {
		int32_t top = heap[0];
		swap_at(0, heap.size() - 1);
		heap.pop_back();
		pos[top] = -1;
		if(!heap.empty())
			down(0);
		return top;
	}
};

uint64_t edge_key(uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; }

} // namespace

extern "C" int lucid_generate_quads(const float *positions, int64_t num_verts, const uint32_t *tris,
									int64_t num_tris, float square_weight, uint32_t *quads_out,
									int64_t *num_quads) {
	if(!positions || !tris || !num_quads || num_tris < 0 || num_verts < 0 ||
	   num_tris >= ((int64_t)1 << 31))
		return LUCID_ERR_INVALID_ARG;
	const size_t nt = (size_t)num_tris;
	for(size_t i = 0; i < 3 * nt; i++)
		if((int64_t)tris[i] >= num_verts)
			return LUCID_ERR_INVALID_ARG;
	// neighbours across each edge (first triangle of a directed edge owns it)
	std::unordered_map<uint64_t, uint32_t> edges;
	edges.reserve(nt * 4);
	for(size_t i = 0; i < nt; i++)
		for(int j = 0; j < 3; j++)
			edges.emplace(edge_key(tris[3 * i + j], tris[3 * i + (j + 1) % 3]), (uint32_t)i);
	std::vector<int32_t> nbr(3 * nt, -1);
	for(size_t i = 0; i < nt; i++)
		for(int j = 0; j < 3; j++) {
			auto it = edges.find(edge_key(tris[3 * i + (j + 1) % 3], tris[3 * i + j]));
			if(it != edges.end() && it->second != (uint32_t)i)
				nbr[3 * i + j] = (int32_t)it->second;
		}
	edges.clear();

	// candidate quads, one per shared edge
	std::vector<Candidate> cand;
	cand.reserve(nt * 3 / 2);
	std::vector<int32_t> tri_quads(3 * nt, -1);
	for(size_t t0 = 0; t0 < nt; t0++) {
		const uint32_t *a = tris + 3 * t0;
		for(int i = 0; i < 3; i++) {
			const int32_t t1 = nbr[3 * t0 + i];
			if(t1 < 0 || tri_quads[3 * t0 + i] != -1)
				continue;
			int j = -1;
			for(int k = 0; k < 3; k++)
				if(nbr[3 * (size_t)t1 + k] == (int32_t)t0) {
					j = k;
					break;
				}
			if(j < 0)
				continue;
			const uint32_t *b = tris + 3 * (size_t)t1;
			int64_t opposite = -1;
			for(int k = 0; k < 3; k++)
				if(b[k] != a[0] && b[k] != a[1] && b[k] != a[2]) {
					opposite = b[k];
					break;
				}
			if(opposite < 0)
				continue;
			Candidate q;
			q.tris[0] = (uint32_t)t0, q.tris[1] = (uint32_t)t1;
			q.verts[0] = a[i], q.verts[1] = (uint32_t)opposite;
			q.verts[2] = a[(i + 1) % 3], q.verts[3] = a[(i + 2) % 3];
			V3f pts[4];
			for(int k = 0; k < 4; k++)
				pts[k] = load(positions, q.verts[k]);
			q.square = squareness(pts);
			const int32_t id = (int32_t)cand.size();
			tri_quads[3 * t0 + i] = id;
			tri_quads[3 * (size_t)t1 + j] = id;
			cand.push_back(q);
		}
	}
	for(size_t t = 0; t < nt; t++)
		for(int i = 0; i < 3; i++) {
			const int32_t q0 = tri_quads[3 * t + i], q1 = tri_quads[3 * t + (i + 1) % 3];
			if(q0 != -1 && q1 != -1) {
				cand[q0].add_conflict(q1);
				cand[q1].add_conflict(q0);
			}
		}

	// greedy independent set
	const size_t nq = cand.size();
	std::vector<uint8_t> state(nq, 0); // 2 selected, 1 removed
	std::vector<uint8_t> tri_done(nt, 0);
	std::vector<int> degree(nq);
	Heap heap(nq);
	auto score = [&](size_t q) { return (float)degree[q] - cand[q].square * square_weight; };
	for(size_t q = 0; q < nq; q++) {
		degree[q] = cand[q].degree();
		heap.insert((int32_t)q, score(q));
	}
	while(!heap.heap.empty()) {
		const int32_t q = heap.pop();
		if(state[q])
			continue;
		state[q] = 2;
		tri_done[cand[q].tris[1]] = 1;
		for(int32_t n : cand[q].conflicts) {
			if(n < 0)
				continue;
			state[n] = 1;
			for(int32_t n2 : cand[n].conflicts)
				if(n2 >= 0 && !state[n2]) {
					degree[n2]--;
					heap.update(n2, score((size_t)n2));
				}
		}
	}

	// output in triangle order
	int64_t count = 0;
	for(size_t t = 0; t < nt; t++) {
		if(tri_done[t])
			continue;
		int32_t sel = -1;
		for(int i = 0; i < 3; i++) {
			const int32_t q = tri_quads[3 * t + i];
			if(q != -1 && state[q] == 2) {
				sel = q;
				break;
			}
		}
		if(quads_out) {
			uint32_t *o = quads_out + 4 * count;
			if(sel < 0)
				o[0] = tris[3 * t], o[1] = tris[3 * t + 1], o[2] = o[3] = tris[3 * t + 2];
			else
				for(int k = 0; k < 4; k++)
					o[k] = cand[sel].verts[k];
		}
		count++;
	}
	*num_quads = count;
	return LUCID_OK;
}
