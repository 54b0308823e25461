// CPU emulation of the patch kernel's indexing (cudafemproject_amd/csrc/
// kernels.hip, k_ebe_patch phases A-C) and of the interface gather over a
// mesh plan built by build_plan: every index the kernels form is checked
// against its buffer, every sorted entry position is written exactly once,
// and the plan-based assembly of random element forces must equal direct
// assembly.  Usage: plan_check <dim> <n> <elems_per_patch> [nodes_per_patch]
#include "femgpu_internal.h"
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <vector>
#include <random>
using namespace femgpu;
#define CHK(c) do { if (!(c)) { printf("FAIL %s line %d\n", #c, __LINE__); exit(1);} } while(0)
int main(int argc, char** argv) {
  int dim = atoi(argv[1]), n = atoi(argv[2]), emax = atoi(argv[3]);
  int pn_max = argc > 4 ? atoi(argv[4]) : 1024;
  std::vector<double> nodes; std::vector<int32_t> elems;
  if (dim == 3) {
    int N = n+1;
    for (int k=0;k<N;++k) for(int j=0;j<N;++j) for(int i=0;i<N;++i){nodes.push_back(i);nodes.push_back(j);nodes.push_back(k);}
    int perms[6][3]={{0,1,2},{0,2,1},{1,0,2},{1,2,0},{2,0,1},{2,1,0}}; int st[3]={1,N,N*N};
    for (int k=0;k<n;++k) for(int j=0;j<n;++j) for(int i=0;i<n;++i) for (auto&p:perms){int v0=i+N*(j+N*k);int v1=v0+st[p[0]];int v2=v1+st[p[1]];int v3=v2+st[p[2]];elems.insert(elems.end(),{v0,v1,v2,v3});}
  } else {
    int nx=n, ny=n/2+1, N=nx+1;
    for (int j=0;j<=ny;++j) for(int i=0;i<=nx;++i){nodes.push_back(i);nodes.push_back(j);}
    for (int j=0;j<ny;++j) for(int i=0;i<nx;++i){int a=i+N*j; elems.insert(elems.end(),{a,a+1,a+N+1}); elems.insert(elems.end(),{a,a+N+1,a+N});}
  }
  int npe=dim+1; long nn=nodes.size()/dim, ne=elems.size()/npe;
  MeshPlan P; std::string m = build_plan(dim, nn, nodes.data(), ne, elems.data(), emax, pn_max, true, P);
  std::vector<uint32_t> rec; pack_records(P, rec); CHK(rec.size() == size_t(ne) * 3);
  // the patch kernel reads record p*emax_used + i for every i < emax_used of
  // patch p: in bounds, equal to the patch's records, zero past them
  { std::vector<uint32_t> r2 = rec, rs; bool reg = stride_records(P, r2, rs); const std::vector<uint32_t>& A = reg ? r2 : rs;
    size_t em = size_t(P.emax_used);
    for (size_t p=0;p<P.patches.size();++p) for (size_t i=0;i<em;++i) { size_t w = (p*em + i)*3; CHK(w + 2 < A.size());
      for (int j=0;j<3;++j) { if (long(i) < P.patches[p].n_elem) CHK(A[w+j] == rec[(size_t(P.patches[p].elem_start)+i)*3+j]); else if (!reg || p + 1 == P.patches.size()) CHK(A[w+j] == 0u); } }
    printf("records %s\n", reg ? "regular" : "strided copy"); }
  if(!m.empty()){printf("plan error: %s\n", m.c_str()); return 1;}
  printf("nn=%ld ne=%ld np=%zu owned=%ld iface=%ld slots=%ld pn_used=%d emax_used=%d\n", nn, ne, P.patches.size(), (long)P.n_owned, (long)P.n_iface, (long)P.n_slots, P.pn_used, P.emax_used);
  // random element "forces": f[e][a][c] ; direct assembly vs plan assembly
  std::mt19937 g(1); std::uniform_real_distribution<double> U(-1,1);
  std::vector<double> fe(ne*npe*dim); for (auto& v: fe) v = U(g);
  // fe indexed by USER element id
  std::vector<double> ydir(nn*dim, 0.0);
  for (long e=0;e<ne;++e) for(int a=0;a<npe;++a) for(int c=0;c<dim;++c) ydir[elems[e*npe+a]*dim+c]+=fe[(e*npe+a)*dim+c];
  std::vector<double> y(nn*dim, 0.0), part(P.n_slots*dim + 1, NAN);
  std::vector<int> slot_written(P.n_slots, 0);
  // the per-patch node tables the matvec kernel reads at stride pn_max (its
  // workgroup size): thread k of patch p reads entry p*pn_max+k (pgid, pend:
  // every k below the largest patch's node count; pslot: interface nodes)
  std::vector<int32_t> pgid, pslot; std::vector<uint16_t> pend;
  pack_patch_nodes(P, pn_max, pgid, pslot, pend);
  CHK(pgid.size() == P.patches.size() * size_t(pn_max) && pslot.size() == pgid.size() && pend.size() == pgid.size());
  CHK(P.node_end.size() == (size_t)std::accumulate(P.patches.begin(), P.patches.end(), 0L, [](long s, const PatchInfo& q){return s + q.n_own + q.n_sh;}));
  for (size_t p=0;p<P.patches.size();++p) {
    const PatchInfo& Q = P.patches[p];
    int nloc = Q.n_own + Q.n_sh; CHK(nloc <= P.pn_used && Q.n_elem <= P.emax_used);
    std::vector<long> gidl(nloc);
    CHK(nloc <= pn_max);
    for (int k=0;k<nloc;++k){ long gid = k<Q.n_own ? Q.own_start+k : P.sh_gid.at(Q.sh_start+k-Q.n_own); CHK(gid>=0 && gid<nn); gidl[k]=gid; CHK(Q.pnode_start + k < (long)P.node_end.size());
      size_t ti = p*size_t(pn_max)+k; CHK(pgid[ti] == gid); CHK(pend[ti] == P.node_end[Q.pnode_start+k]);
      if (k<Q.n_own) CHK(pslot[ti] == 0);
      else { uint32_t w = uint32_t(pslot[ti]); long slot = long(w & 0x7fffffffu); CHK(slot == P.sh_slot[Q.sh_start+k-Q.n_own]);
             CHK(((w >> 31) != 0) == (slot == P.slot_start[gid-P.n_owned])); } }
    for (int k=nloc;k<pn_max;++k){ size_t ti = p*size_t(pn_max)+k; CHK(pgid[ti] == 0 && pslot[ti] == 0 && pend[ti] == 0); }
    // the kernel zeroes its entry list first; padding entries stay zero
    CHK(P.ent_used <= 16383); std::vector<double> res(P.ent_used*dim, 0.0); std::vector<int> written(P.ent_used, 0);
    for (int i=0;i<Q.n_elem;++i){ long ie = Q.elem_start+i; CHK(ie < ne); long ue = P.elem_perm[ie];
      // the kernel's decoding of the 12-byte record
      uint32_t w0 = rec[ie*3], w1 = rec[ie*3+1], w2 = rec[ie*3+2];
      int L[4] = {int(w0 & 1023u), int((w0 >> 10) & 1023u), int((w0 >> 20) & 1023u), int((w0 >> 30) | ((w1 & 255u) << 2))};
      int Pq[4] = {int((w1 >> 8) & 16383u), int((w1 >> 22) | ((w2 & 15u) << 10)), int((w2 >> 4) & 16383u), int(w2 >> 18)};
      for (int a=0;a<npe;++a){ int l=L[a]; CHK(l == P.elem_lid[ie*4+a]); CHK(l<nloc); CHK(P.node_perm[gidl[l]] == elems[ue*npe+a]); CHK(P.elems_int[ie*npe+a]==gidl[l]);
        int pos = Pq[a]; CHK(pos == P.epos[ie*4+a]); CHK(pos >= 0 && pos < P.ent_used); CHK(!written[pos]); written[pos]=1;
        { int bb0 = l? P.node_end[Q.pnode_start+l-1]:0, bb1 = P.node_end[Q.pnode_start+l]; CHK(pos >= bb0 && pos < bb1); }
        for(int c=0;c<dim;++c) res[pos*dim+c] = fe[(ue*npe+a)*dim+c]; } }
    for (int k=0;k<nloc;++k){ size_t ti = p*size_t(pn_max)+k; int b0 = k? pend[ti-1]:0, b1=pend[ti]; CHK(b0<=b1 && b1 <= P.ent_used); CHK(b0 % 4 == 0 && b1 % 4 == 0);
      { int nw = 0; for (int j=b0;j<b1;++j) nw += written[j]; CHK(nw > 0 && b1 - b0 - nw < 4); }
      double acc[3]={0,0,0};
      for (int j=b0;j<b1;++j){ for(int c=0;c<dim;++c){ double v=res[j*dim+c]; acc[c]+=v; } }
      if (k<Q.n_own){ long gid=pgid[ti]; CHK(gid < P.n_owned); for(int c=0;c<dim;++c) y[gid*dim+c]=acc[c]; }
      else { long slot=long(uint32_t(pslot[ti]) & 0x7fffffffu); CHK(slot>=0 && slot<P.n_slots); CHK(!slot_written[slot]); slot_written[slot]=1; for(int c=0;c<dim;++c) part[slot*dim+c]=acc[c]; } }
  }
  for (long s=0;s<P.n_iface;++s){ double acc[3]={0,0,0}; for(int j=P.slot_start[s]; j<P.slot_start[s+1]; ++j){ CHK(slot_written[j]); for(int c=0;c<dim;++c) acc[c]+=part[j*dim+c]; } for(int c=0;c<dim;++c) y[(P.n_owned+s)*dim+c]=acc[c]; }
  double err=0; for (long i=0;i<nn;++i) for(int c=0;c<dim;++c) err=std::max(err, std::fabs(y[i*dim+c]-ydir[P.node_perm[i]*dim+c]));
  printf("max err %.3e\n", err); CHK(err < 1e-12);
  return 0;
}
