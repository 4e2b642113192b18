// kernels.hip - gfx950 kernels of the wavefront path tracer.
//
// One RtRender call runs, per chunk of R sample groups (R*npix path slots):
//   reset counters -> raygen -> for each bounce: closest-hit traversal (persistent lanes)
//   -> shade (NEE shadow request + next ray, compacted with wave-aggregated atomics)
//   -> any-hit traversal of the shadow queue (adds the NEE term when unoccluded)
//   -> accumulate (adds the R groups to the FP32 sums in sample order, tone-maps).
// Queue sizes live in device memory, so no host synchronisation happens inside a frame.
#include "kernels.h"
#include "traverse.h"

namespace rt {

__device__ const unsigned gSobolDirs[kSobolDims * 32] = {
#include "sobol_directions.inc"
};
static const unsigned kSobolDirsHost[kSobolDims * 32] = {
#include "sobol_directions.inc"
};
const unsigned* hostSobolDirections() { return kSobolDirsHost; }

// ------------------------------------------------------------------ traversal
template <int W>
__device__ __forceinline__ void loadNode(const WideNode<W>* nodes, int ref, WideNode<W>& n) {
    union { float4 v[2 * W]; WideNode<W> node; } u;
    const float4* p = reinterpret_cast<const float4*>(nodes + ref);
#pragma unroll
    for (int j = 0; j < 2 * W; j++) u.v[j] = p[j];
    n = u.node;
}

template <int W, bool ANY, bool STATS>
__global__ __launch_bounds__(kBlock) void traceKernel(TraceArgs a) {
    // Short stack: kShortStack (ref, entry distance) pairs per lane in LDS, entry-major so
    // the 64 lanes of a wave touch 64 consecutive words; deeper entries go to a global ring.
    __shared__ int sRef[kShortStack * kBlock];
    __shared__ float sDist[kShortStack * kBlock];
    const int tid = threadIdx.x, lane = tid & 63;
    const int gl = blockIdx.x * kBlock + tid;
    const int total = a.countPtr ? *a.countPtr : a.count;
    const WideNode<W>* nodes = static_cast<const WideNode<W>*>(a.nodes);
    const float4* recs = reinterpret_cast<const float4*>(a.recs);
    int2* spill = reinterpret_cast<int2*>(a.spill);
    unsigned long long nv = 0, tt = 0;
    bool overflow = false;
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(a.work, 64);
        base = __shfl(base, 0, 64);
        if (base >= total) break;
        const int i = base + lane;
        if (i >= total) continue;
        const float4 r0 = a.rays[2 * i], r1 = a.rays[2 * i + 1];
        const f3 o = F3(r0.x, r0.y, r0.z), d = F3(r1.x, r1.y, r1.z);
        const f3 inv = safeInverse(d);
        float best = r0.w, bu = 0, bv = 0;
        int bestIdx = -1;
        int sp = 0, ref = 0;
        bool done = false;
        for (;;) {
            if (ref >= 0) {
                WideNode<W> n;
                loadNode<W>(nodes, ref, n);
                if (STATS) nv++;
                float dist[W]; int cr[W];
#pragma unroll
                for (int c = 0; c < W; c++) {
                    dist[c] = boxEntry(o, inv, n.lox[c], n.loy[c], n.loz[c], n.hix[c], n.hiy[c], n.hiz[c], best);
                    cr[c] = n.child[c];
                }
                sortChildren<W>(dist, cr);
#pragma unroll
                for (int c = W - 1; c >= 1; c--) {
                    if (dist[c] != INFINITY) {
                        const int slot = sp & (kShortStack - 1);
                        if (sp >= kShortStack) {
                            const int e = sp - kShortStack;
                            if (e < a.spillEntries)
                                spill[(size_t)e * a.totalLanes + gl] = make_int2(sRef[slot * kBlock + tid], __float_as_int(sDist[slot * kBlock + tid]));
                            else
                                overflow = true;
                        }
                        sRef[slot * kBlock + tid] = cr[c];
                        sDist[slot * kBlock + tid] = dist[c];
                        sp++;
                    }
                }
                if (dist[0] != INFINITY) { ref = cr[0]; continue; }
            } else {
                for (int k = ~ref;; k++) {
                    const float4 A = recs[3 * k], B = recs[3 * k + 1], C = recs[3 * k + 2];
                    const int flags = __float_as_int(B.w), index = __float_as_int(A.w);
                    float u = 0, v = 0;
                    const float t = intersectTri(o, d, F3(A.x, A.y, A.z), F3(B.x, B.y, B.z), F3(C.x, C.y, C.z),
                                                 (flags & kTriOneSided) != 0, u, v);
                    if (STATS) tt++;
                    if (t > 0 && (t < best || (t == best && index < bestIdx))) {
                        best = t; bestIdx = index; bu = u; bv = v;
                        if (ANY) { done = true; break; }
                    }
                    if (flags & kTriLastInLeaf) break;
                }
                if (done) break;
            }
            bool found = false;
            while (sp > 0) {
                sp--;
                const int slot = sp & (kShortStack - 1);
                const int eref = sRef[slot * kBlock + tid];
                const float edist = sDist[slot * kBlock + tid];
                if (sp >= kShortStack) {
                    const int e = sp - kShortStack;
                    if (e < a.spillEntries) {
                        const int2 back = spill[(size_t)e * a.totalLanes + gl];
                        sRef[slot * kBlock + tid] = back.x;
                        sDist[slot * kBlock + tid] = __int_as_float(back.y);
                    }
                }
                if (edist <= best) { ref = eref; found = true; break; }
            }
            if (!found) break;
        }
        if (a.hits) {
            a.hits[i] = make_int4(bestIdx, __float_as_int(best), __float_as_int(bu), __float_as_int(bv));
        } else if (bestIdx < 0) {
            // Shadow query found no occluder: add the stored next-event term.
            const int slot = __float_as_int(r1.w);
            float4 acc = a.rad[slot];
            const float4 c = a.contrib[slot];
            acc.x = acc.x + c.x; acc.y = acc.y + c.y; acc.z = acc.z + c.z;
            a.rad[slot] = acc;
        }
    }
    if (STATS) {
        atomicAdd(&a.stats[0], nv);
        atomicAdd(&a.stats[1], tt);
    }
    if (overflow) atomicOr(a.overflow, 1);
}

template <int W>
static hipError_t launchTraceW(bool anyHit, bool stats, int blocks, const TraceArgs& a, hipStream_t s) {
    dim3 grid(blocks), block(kBlock);
    if (anyHit) {
        if (stats) hipLaunchKernelGGL((traceKernel<W, true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((traceKernel<W, true, false>), grid, block, 0, s, a);
    } else {
        if (stats) hipLaunchKernelGGL((traceKernel<W, false, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((traceKernel<W, false, false>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}
hipError_t launchTrace(int width, bool anyHit, bool stats, int blocks, const TraceArgs& a, hipStream_t s) {
    switch (width) {
        case 2: return launchTraceW<2>(anyHit, stats, blocks, a, s);
        case 4: return launchTraceW<4>(anyHit, stats, blocks, a, s);
        case 8: return launchTraceW<8>(anyHit, stats, blocks, a, s);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------ frame kernels
__global__ void resetCountersKernel(int* c, int initialRays) {
    for (int k = threadIdx.x; k < kCounterStride; k += blockDim.x) c[k] = 0;
    __syncthreads();
    if (threadIdx.x == 0) c[kCntRays] = initialRays;
}
hipError_t launchResetCounters(int* counters, int initialRays, hipStream_t s) {
    hipLaunchKernelGGL(resetCountersKernel, dim3(1), dim3(kBlock), 0, s, counters, initialRays);
    return hipGetLastError();
}

// Sobol terms shared by a launch: shift[k] depends on (seed, dim), term[r][k] on
// (sample, dim); a lane combines them with its pixel hash (Native/sampling.cuh:26-35).
template <int K>
__device__ __forceinline__ void loadSobolTerms(unsigned* sShift, unsigned* sTerm, unsigned seed, unsigned dimBase,
                                               unsigned sample0, int R) {
    const int tid = threadIdx.x;
    if (tid < K) sShift[tid] = dimShift(seed, dimBase + tid);
    if (tid < R * K) sTerm[tid] = grayTerm(gSobolDirs, dimBase + (tid % K), sample0 + tid / K);
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void raygenKernel(FrameArgs f) {
    __shared__ unsigned sShift[4], sTerm[kMaxRaysPerPixel * 4];
    loadSobolTerms<4>(sShift, sTerm, f.rs.seed, 0, f.sample0, f.R);
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot >= f.R * f.npix) return;
    const int r = slot / f.npix;
    const unsigned pixel = (unsigned)(slot - r * f.npix);
    const int x = (int)(pixel % (unsigned)f.w), y = (int)(pixel / (unsigned)f.w);
    auto rng = [&](int k) { return sobolFromParts(pixel, sShift[k], sTerm[r * 4 + k]); };
    f3 o, d;
    cameraRay(f.cam, f.rs, x, y, f.w, f.h, rng, o, d);
    f.thr[slot] = make_float4(1, 1, 1, 0);
    f.rad[slot] = make_float4(0, 0, 0, 0);
    f.alb[slot] = make_float4(0, 0, 0, 0);
    f.nrm[slot] = make_float4(0, 0, 0, 0);
    f.rayOut[2 * slot] = make_float4(o.x, o.y, o.z, 1e30f);
    f.rayOut[2 * slot + 1] = make_float4(d.x, d.y, d.z, __int_as_float(slot));
}
hipError_t launchRaygen(const FrameArgs& f, hipStream_t s) {
    const int n = f.R * f.npix;
    hipLaunchKernelGGL(raygenKernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, f);
    return hipGetLastError();
}

// Appends one 32-byte ray per predicated lane with one atomic per wave.
__device__ __forceinline__ void appendRay(bool pred, int* counter, float4* queue, float4 a, float4 b) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader, 64);
    if (pred) {
        const int off = base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
        queue[2 * off] = a;
        queue[2 * off + 1] = b;
    }
}

__global__ __launch_bounds__(kBlock) void shadeKernel(FrameArgs f) {
    __shared__ unsigned sShift[8], sTerm[kMaxRaysPerPixel * 8];
    loadSobolTerms<8>(sShift, sTerm, f.rs.seed, 4 + 8 * f.bounce, f.sample0, f.R);
    const int count = f.counters[kCntRays + f.bounce];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (blockIdx.x * kBlock >= count) return;   // whole block idle (uniform exit)
    bool wantShadow = false, wantNext = false;
    float4 sa = make_float4(0, 0, 0, 0), sb = sa, na = sa, nb = sa;
    if (i < count) {
        const float4 r0 = f.rayIn[2 * i], r1 = f.rayIn[2 * i + 1];
        const int slot = __float_as_int(r1.w);
        const int r = slot / f.npix;
        const unsigned pixel = (unsigned)(slot - r * f.npix);
        const int4 hv = f.hits[i];
        RtHit hit;
        hit.index = hv.x; hit.distance = __int_as_float(hv.y); hit.u = __int_as_float(hv.z); hit.v = __int_as_float(hv.w);
        const float4 t4 = f.thr[slot], q4 = f.rad[slot];
        PathState ps;
        ps.thr = F3(t4.x, t4.y, t4.z); ps.prevPdf = t4.w; ps.rad = F3(q4.x, q4.y, q4.z);
        FirstHit fh; fh.covered = false;
        auto rng = [&](int k) { return sobolFromParts(pixel, sShift[k], sTerm[r * 8 + k]); };
        ShadeResult res;
        shadeBounce(f.sc, f.rs, f.bounce, rng, F3(r0.x, r0.y, r0.z), F3(r1.x, r1.y, r1.z), hit, ps, &fh, res);
        f.thr[slot] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, ps.prevPdf);
        f.rad[slot] = make_float4(ps.rad.x, ps.rad.y, ps.rad.z, q4.w);
        if (fh.covered) {
            f.alb[slot] = make_float4(fh.albedo.x, fh.albedo.y, fh.albedo.z, 1.0f);
            f.nrm[slot] = make_float4(fh.normal.x, fh.normal.y, fh.normal.z, fh.depth);
        }
        if (res.shadow) {
            wantShadow = true;
            f.nee[slot] = make_float4(res.contrib.x, res.contrib.y, res.contrib.z, 0);
            sa = make_float4(res.nextO.x, res.nextO.y, res.nextO.z, res.slimit);
            sb = make_float4(res.sdir.x, res.sdir.y, res.sdir.z, __int_as_float(slot));
        }
        if (res.cont) {
            wantNext = true;
            na = make_float4(res.nextO.x, res.nextO.y, res.nextO.z, 1e30f);
            nb = make_float4(res.nextD.x, res.nextD.y, res.nextD.z, __int_as_float(slot));
        }
    }
    appendRay(wantShadow, &f.counters[kCntShadow + f.bounce], f.shadow, sa, sb);
    appendRay(wantNext, &f.counters[kCntRays + f.bounce + 1], f.rayOut, na, nb);
}
hipError_t launchShade(const FrameArgs& f, hipStream_t s) {
    const int n = f.R * f.npix;
    hipLaunchKernelGGL(shadeKernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, f);
    return hipGetLastError();
}

// Native/renderer.cu:292-301: add the R paths of each pixel in sample order, then tone-map.
__global__ __launch_bounds__(kBlock) void accumulateKernel(AccumArgs a) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.npix) return;
    float4 sr = make_float4(0, 0, 0, 0), sa = sr, sn = sr;
    if (!a.startFromZero) { sr = a.sumRad[i]; sa = a.sumAlb[i]; sn = a.sumNrm[i]; }
    for (int r = 0; r < a.R; r++) {
        const size_t slot = (size_t)r * a.npix + i;
        const float4 pr = a.rad[slot], pa = a.alb[slot], pn = a.nrm[slot];
        sr.x = sr.x + pr.x; sr.y = sr.y + pr.y; sr.z = sr.z + pr.z;
        sa.x = sa.x + pa.x; sa.y = sa.y + pa.y; sa.z = sa.z + pa.z;
        sn.x = sn.x + pn.x; sn.y = sn.y + pn.y; sn.z = sn.z + pn.z;
        sr.w = sr.w + pa.w;   // covered count
        sn.w = sn.w + pn.w;   // first-hit depth sum
    }
    a.sumRad[i] = sr; a.sumAlb[i] = sa; a.sumNrm[i] = sn;
    if (a.writeDisplay) {
        const float e = a.exposure;
        a.pixels[i] = make_uchar4(displayByte(sr.x * a.invSamples, e), displayByte(sr.y * a.invSamples, e),
                                  displayByte(sr.z * a.invSamples, e), 255);
    }
}
hipError_t launchAccumulate(const AccumArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(accumulateKernel, dim3((a.npix + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ Sobol batch
__global__ __launch_bounds__(kBlock) void sobolBatchKernel(const unsigned* pixels, const unsigned* samples, int count,
                                                           unsigned seed, int dims, float* out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const unsigned p = pixels[i], s = samples[i];
    for (int d = 0; d < dims; d++) out[(size_t)i * dims + d] = sobolSample(gSobolDirs, p, s, seed, (unsigned)d);
}
hipError_t launchSobolBatch(const unsigned* pixels, const unsigned* samples, int count, unsigned seed, int dims,
                            float* out, hipStream_t s) {
    hipLaunchKernelGGL(sobolBatchKernel, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pixels, samples,
                       count, seed, dims, out);
    return hipGetLastError();
}

}  // namespace rt
