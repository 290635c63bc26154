// txg_fused.h -- one-pass D3Q19 step: a workgroup owns an xy-tile of
// TX x TY nodes and marches along a z-segment.
//
// The psi-gradient force at plane z needs the density of planes z-r..z+r
// on the tile plus an r-wide ring (r = 1 at isotropy 4, 2 at isotropy 8),
// and that density needs the pulled populations of those nodes.  Per plane
// the workgroup
//   A. pull-sums the density of plane z+r over tile+ring into an LDS ring
//      buffer of 2r+1 planes (NaN marks walls and cells beyond non-periodic
//      faces),
//   B. re-pulls its TX x TY interior nodes of plane z (read r planes ago by
//      the same workgroup, so served from L2/MALL) and runs moments, forces
//      (stencil density from LDS), common velocity and collision, writing f*.
// HBM therefore sees the populations read once and written once per step,
// versus read twice for the two-pass step.  Without a fluid-fluid force
// (r = 0) phase A disappears and the density is the node's own.
//
// Populations are addressed through one buffer resource per component
// (base in SGPRs, 32-bit byte offsets per lane, hardware range check), so a
// pull costs a few 32-bit adds and a select instead of 64-bit address
// arithmetic.  This requires Q*Np*sizeof(S) < 4 GiB per component, which
// the host checks before choosing this kernel.
//
// Work decomposition: tiles x z-segments, with the workgroup index remapped
// so that consecutive tiles of a segment land on the same XCD (the hardware
// hands workgroups to the 8 XCDs round-robin), keeping ring and re-pull
// traffic within one L2.
#pragma once

namespace txg {

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef __amdgpu_buffer_rsrc_t rsrc_t;

// buffer resource over `bytes` bytes at p (raw addressing, range-checked)
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

template <class S, class R>
__device__ __forceinline__ R bld(rsrc_t r, uint32_t off, R shift) {
  if constexpr (std::is_same<S, float>::value) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
  } else if constexpr (std::is_same<S, double>::value) {
    u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
    return __builtin_bit_cast(double, v);
  } else {
    unsigned short u = __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0);
    return (R)__half2float(__ushort_as_half(u)) + shift;
  }
}

template <class S, class R>
__device__ __forceinline__ void bst(rsrc_t r, uint32_t off, R v, R shift) {
  if constexpr (std::is_same<S, float>::value) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)off, 0, 0);
  } else if constexpr (std::is_same<S, double>::value) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)off, 0, 0);
  } else {
    __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(__float2half_rn((float)(v - shift))), r, (int)off, 0, 0);
  }
}

// byte offsets of the +-1 rows of a column: x, y*NX, z*NX*NY times sizeof(S)
struct RowsB {
  uint32_t x[3], y[3], z[3];
};

// pull-stream with bounce-back (DistributionStream + DistributionBounceback,
// lbm_distribution_function.F90:560-784) through buffer resources
template <int C, class S, class R>
__device__ __forceinline__ void pull_buf(const Params<R>& p, const rsrc_t (&rin)[C], uint32_t mk, const RowsB& rb,
                                         uint32_t plane_b, R (&f)[C][19]) {
  using L = Lat<3>;
  const uint32_t node_b = rb.x[1] + rb.y[1] + rb.z[1];
  sfor<0, 19>([&](auto NI) {
    constexpr int n = decltype(NI)::value;
    uint32_t off;
    if constexpr (n == 0) {
      off = node_b;
    } else {
      const uint32_t so = rb.x[1 - L::cx[n]] + rb.y[1 - L::cy[n]] + rb.z[1 - L::cz[n]] + (uint32_t)n * plane_b;
      const uint32_t bo = node_b + (uint32_t)L::opp[n] * plane_b;
      off = (mk & (1u << n)) ? bo : so;
    }
#pragma unroll
    for (int m = 0; m < C; m++) f[m][n] = bld<S, R>(rin[m], off, (R)L::w[n] * p.shift0[m]);
  });
}

template <class S, class R>
struct BufStore {
  const Params<R>& p;
  const rsrc_t* rout;
  uint32_t node_b, plane_b;
  __device__ __forceinline__ void operator()(int m, int n, R v) const {
    bst<S, R>(rout[m], node_b + (uint32_t)n * plane_b, v, (R)Lat<3>::w[n] * p.shift0[m]);
  }
};

// LDS stencil density: s[(slot*C + m)*EY*EX + row*EX + col]
template <class R, int C, int RR, int EXW, int EYW>
struct LdsRho {
  const R* s;
  int col;                 // linear column of the node in the extended tile
  int slot[2 * RR + 1];    // ring-buffer slot of plane z + dz, dz = -RR..RR
  template <int X, int Y, int Z>
  __device__ __forceinline__ R get(int m) const {
    return s[(slot[Z + RR] * C + m) * (EYW * EXW) + col + Y * EXW + X];
  }
  template <int X, int Y, int Z>
  __device__ __forceinline__ bool fluid() const {
    return is_num(get<X, Y, Z>(0));
  }
};

struct FusedGeom {
  int ntx, nty, ntiles, nseg, zseg, nblocks;
};

// map the hardware workgroup index to a logical one so that each XCD gets a
// contiguous range of logical workgroups
__device__ __forceinline__ int xcd_remap(int i, int nblocks) {
  constexpr int NX = 8;
  const int per = nblocks / NX, rem = nblocks % NX;
  const int xcd = i % NX, k = i / NX;
  return xcd * per + min(xcd, rem) + k;
}

template <int C, int NFF, bool FAST, class S, class R, int TX, int TY>
__global__ void __launch_bounds__(TX* TY) k_fused3(Params<R> p, Bufs<S, R> b, FusedGeom g) {
  using L = Lat<3>;
  constexpr int Q = L::Q;
  constexpr int RR = NFF == kFF3Count8 ? 2 : (NFF == kFF3Count4 ? 1 : 0);
  constexpr int EXW = TX + 2 * RR, EYW = TY + 2 * RR, NCOL = EXW * EYW, NPL = 2 * RR + 1;
  constexpr int NT = TX * TY;
  constexpr int NCP = (NCOL + NT - 1) / NT;  // columns per thread in phase A
  constexpr uint32_t SB = sizeof(S);
  __shared__ R srho[RR > 0 ? NPL * C * NCOL : 1];

  const uint32_t plane_b = (uint32_t)p.Np * SB;
  rsrc_t rin[C], rout[C];
#pragma unroll
  for (int m = 0; m < C; m++) {
    rin[m] = make_rsrc(b.fin + (size_t)m * Q * p.Np, (uint32_t)Q * plane_b);
    rout[m] = make_rsrc(b.fout + (size_t)m * Q * p.Np, (uint32_t)Q * plane_b);
  }

  const int lb = xcd_remap(blockIdx.x, g.nblocks);
  const int tile = lb % g.ntiles, seg = lb / g.ntiles;
  const int x0 = (tile % g.ntx) * TX, y0 = (tile / g.ntx) * TY;
  const int z0 = seg * g.zseg, z1 = min(p.nz, z0 + g.zseg);
  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
  const int x = x0 + tx, y = y0 + ty;
  const bool active = x < p.nx && y < p.ny;

  // rows of the interior column; nb (element units) serves the wettability force
  RowsB rb;
  NbrH<1> nb;
  if (active) {
    make_nbr<3, 1, R>(p, x, y, z0, nb);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      rb.x[d] = (uint32_t)nb.xw[d] * SB;
      rb.y[d] = (uint32_t)nb.yw[d] * SB;
    }
  }

  // phase-A columns of this thread
  uint32_t cx[NCP][3], cy[NCP][3];
  bool cval[NCP];
#pragma unroll
  for (int k = 0; k < NCP; k++) {
    const int c = t + k * NT;
    cval[k] = false;
    if (RR > 0 && c < NCOL) {
      const int gx = wrap_coord(x0 - RR + c % EXW, p.nx, p.periodic[0]);
      const int gy = wrap_coord(y0 - RR + c / EXW, p.ny, p.periodic[1]);
      cval[k] = gx >= 0 && gy >= 0;
#pragma unroll
      for (int d = -1; d <= 1; d++) {
        const int xx = gx < 0 ? -1 : wrap_coord(gx + d, p.nx, p.periodic[0]);
        const int yy = gy < 0 ? -1 : wrap_coord(gy + d, p.ny, p.periodic[1]);
        cx[k][d + 1] = (uint32_t)xx * SB;
        cy[k][d + 1] = (uint32_t)(yy * p.nx) * SB;
      }
    }
  }

  // phase A: density of plane zz over tile + ring into slot
  auto phase_a = [&](int zz, int slot) {
    const int zc = wrap_coord(zz, p.nz, p.periodic[2]);
    RowsB cr;
#pragma unroll
    for (int d = -1; d <= 1; d++) {
      const int w = zc < 0 ? -1 : wrap_coord(zc + d, p.nz, p.periodic[2]);
      cr.z[d + 1] = (uint32_t)(w * p.nxy) * SB;
    }
#pragma unroll
    for (int k = 0; k < NCP; k++) {
      const int c = t + k * NT;
      if (c >= NCOL) continue;
      R rho[C];
#pragma unroll
      for (int m = 0; m < C; m++) rho[m] = (R)NAN;
      if (cval[k] && zc >= 0) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
          cr.x[d] = cx[k][d];
          cr.y[d] = cy[k][d];
        }
        const uint32_t node = (cr.x[1] + cr.y[1] + cr.z[1]) / SB;
        const uint32_t mk = b.mask[node];
        if (!(mk & kSelfSolid)) {
          R f[C][Q];
          pull_buf<C, S, R>(p, rin, mk, cr, plane_b, f);
#pragma unroll
          for (int m = 0; m < C; m++) {
            R s = (R)0;
#pragma unroll
            for (int n = 0; n < Q; n++) s = s + f[m][n];
            rho[m] = s;
          }
          if (mk & (0x3fu << kFaceShift)) {
#pragma unroll
            for (int fc = 0; fc < 6; fc++)
              if (mk & (1u << (kFaceShift + fc)))
#pragma unroll
                for (int m = 0; m < C; m++) rho[m] = p.bcrho[fc][m];
          }
        }
      }
#pragma unroll
      for (int m = 0; m < C; m++) srho[(slot * C + m) * NCOL + c] = rho[m];
    }
  };

  auto slot_of = [&](int zz) { return (zz - (z0 - RR) + NPL * 4096) % NPL; };

  if constexpr (RR > 0) {
    for (int zz = z0 - RR; zz < z0 + RR; zz++) phase_a(zz, slot_of(zz));
  }
  for (int z = z0; z < z1; z++) {
    if constexpr (RR > 0) {
      phase_a(z + RR, slot_of(z + RR));
      __syncthreads();
    }
    if (active) {
      set_z<3, 1, R>(p, z, nb);
#pragma unroll
      for (int d = 0; d < 3; d++) rb.z[d] = (uint32_t)nb.zw[d] * SB;
      const int node = nb.xw[1] + nb.yw[1] + nb.zw[1];
      const uint32_t mk = b.mask[node];
      if (!(mk & kSelfSolid)) {
        R f[C][Q];
        pull_buf<C, S, R>(p, rin, mk, rb, plane_b, f);
        R rs[C];
        if constexpr (RR == 0) {
#pragma unroll
          for (int m = 0; m < C; m++) {
            R s = (R)0;
#pragma unroll
            for (int n = 0; n < Q; n++) s = s + f[m][n];
            rs[m] = s;
          }
          if (mk & (0x3fu << kFaceShift)) {
#pragma unroll
            for (int fc = 0; fc < 6; fc++)
              if (mk & (1u << (kFaceShift + fc)))
#pragma unroll
                for (int m = 0; m < C; m++) rs[m] = p.bcrho[fc][m];
          }
        }
        if (mk & (0x3fu << kFaceShift)) apply_dirichlet<3, C, R>(p, mk, f);
        LdsRho<R, C, RR, EXW, EYW> acc;
        acc.s = srho;
        acc.col = (ty + RR) * EXW + tx + RR;
#pragma unroll
        for (int dz = -RR; dz <= RR; dz++) acc.slot[dz + RR] = slot_of(z + dz);
        if constexpr (RR > 0) {
#pragma unroll
          for (int m = 0; m < C; m++) rs[m] = acc.template get<0, 0, 0>(m);
        }
        R rl[C], j[C][3], F[C][3], up[3];
        moments_forces<3, C, NFF, FAST, R>(p, b.tag, mk, nb, f, rs, acc, rl, j, F, up);
        collide_store<3, C, FAST, R>(p, BufStore<S, R>{p, rout, (uint32_t)node * SB, plane_b}, f, rl, F, up);
      }
    }
    if constexpr (RR > 0) __syncthreads();
  }
}

}  // namespace txg
