// shockx.hip -- libshockx.so: gfx950 kernels and the C ABI of the SHOCK RHS / RK-stage path.
//
// Stage pipeline (one RK stage; src/Functions.c:2107-2365):
//   k_ghost_periodic   on-device ghost fill of rho,u,v,w,e along periodic directions (replaces
//                      the MPI self-interface halo, src/Functions.c:3580-3684)
//   k_calc_values      T and mu on every padded node (CalcValues, src/Functions.c:2392-2481;
//                      p, c and theta are recomputed inside the flux kernels instead of stored)
//   k_face_bc          boundary writers of non-periodic faces, in face order (src/BC.c:13-74)
//   k_visc_grad_march  nodal derivatives of u,v,w,T (src/ZD.c:697-810); k_visc_grad in 2D
//   k_weno             one launch per direction: LDS-staged pencil, characteristic WENO
//                      interface fluxes and their divergence fused, interface fluxes never
//                      reach HBM (CalcFluxesIn*Direction, src/WENO.c:206-1083)
//   k_visc_sweep       one launch per direction: viscous face fluxes and their divergence
//                      (src/ZD.c:47-694); the last one also applies the RK stage update and
//                      the source (src/Functions.c:2287-2365) and writes the next stage state
//                      into another buffer (the reference updates in place after all fluxes of
//                      the stage are assembled, which the separate buffer reproduces exactly)
//   k_rk_update        source + RK update for inviscid runs
// Every hot kernel body is also reachable through a k_*_batch twin that runs many cases from
// one launch over a tile work list (shx_batch_*).
#include "shockx_kernels.h"
#include "../../include/shockx.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <type_traits>
#include <algorithm>
#include <thread>
#include <vector>

using namespace shx;

// ----------------------------------------------------------------------------------------
// kernels
// ----------------------------------------------------------------------------------------

// Periodic ghost fill of nf fields: every node that is a ghost in a periodic direction copies
// the node at the wrapped index of each periodic direction (fixed point of the ghost-inclusive
// halo passes, src/Functions.c:2117-2136; indices along non-periodic directions are kept, so a
// face slab's ghosts wrap along the periodic tangential directions only).
// node index of the periodic image of padded node idx; ghost: idx is a ghost in a periodic direction
__device__ __forceinline__ long periodic_image(const Geom &gm, const int idx[3], bool &ghost)
{
    ghost = false;
    long src = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int s = idx[d];
        if (gm.ng[d] > 1 && gm.per[d]) {
            if (s < gm.lo[d]) { s += gm.n[d]; ghost = true; }
            else if (s >= gm.lo[d] + gm.n[d]) { s -= gm.n[d]; ghost = true; }
        }
        src += (long)s * gm.st[d];
    }
    return src;
}
__device__ __forceinline__ void ghost_body(const Fields10 &f, const int nf, const Geom &gm, const long t)
{
    if (t >= gm.ntot) return;
    int idx[3];
    const long n = padded_node(gm, t, idx);
    bool ghost;
    const long src = periodic_image(gm, idx, ghost);
    if (!ghost) return;
#pragma unroll
    for (int m = 0; m < 10; m++)
        if (m < nf) f.f[m][n] = f.f[m][src];
}
__global__ void k_ghost_periodic(Fields10 f, int nf, Geom gm)
{
    ghost_body(f, nf, gm, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- face boundary writers ---------------------------------------------------------------
// One launch per non-periodic face, in the reference's face order, after CalcValues.  A thread
// owns one ghost node of the face slab: layer l outside the face, tangential indices (i1, i2).
// Its mirror node is layer l inside the face (StartReal + l, EndReal - l).
// Cases with a subsonic outflow face keep the stored p and c of every node (P, C): CalcValues
// values, overwritten by every face writer on its ghosts.  Those writers store p and c that the
// equation of state does not give from rho,u,v,w,e, and the flux kernels read them there.
struct BcArgs {
    Fields5 pr;              // stage input state; ghost rho,u,v,w,e are written
    gdp T, mu;               // derived T and mu (viscous runs; null when inviscid)
    gcdp bs[6];              // exact rho,u,v,w,e,p for manufactured faces
    gcdp met[9];             // metric rows (subsonic outflow writers)
    gdp P, C;                // stored p and c (null unless a face is a subsonic outflow)
    gdp PL, RL, UL;          // Rudy: step-start ghost p, rho, u (written by face_save_body)
    gcdp C0;                 // Rudy: sound speed of the uploaded state
    int kind, dir, upper, g;
    int t1, t2;              // tangential axes
    int lo1, n1, lo2, n2;    // tangential index ranges
    int first, last;         // first / last interior index along dir
    long st[3];
    double rho_in, u_in, v_in, w_in, p_in, cinf, mach;
    double T_wall;           // isothermal walls
    double p_out, alpha_tau; // subsonic outflow: p_out and AlphaNonRef * tau
    Consts k;
};

// ghost node n and mirror node s of work item t of a face launch; false past the slab
__device__ __forceinline__ bool face_node(const BcArgs &a, const long t, long &n, long &s)
{
    const long per_layer = (long)a.n1 * a.n2;
    if (t >= per_layer * a.g) return false;
    const int l = (int)(t / per_layer);
    const long r = t - (long)l * per_layer;
    const int i1 = a.lo1 + (int)(r / a.n2), i2 = a.lo2 + (int)(r % a.n2);
    const int gi = a.upper ? a.last + 1 + l : a.first - 1 - l;
    const int si = a.upper ? a.last - l : a.first + l;
    const long base = (long)i1 * a.st[a.t1] + (long)i2 * a.st[a.t2];
    n = base + (long)gi * a.st[a.dir];
    s = base + (long)si * a.st[a.dir];
    return true;
}

// Rudy faces at the start of a step: U_lastStep's ghost p, rho and u, which the writers of every
// stage read (the first stage's writers overwrite the state buffer's ghosts)
__device__ __forceinline__ void face_save_body(const BcArgs &a, const long t)
{
    long n, s;
    if (!face_node(a, t, n, s)) return;
    a.PL[n] = a.P[n];
    a.RL[n] = a.pr.f[0][n];
    a.UL[n] = a.pr.f[1][n];
}

// contravariant velocities of the mirror node s (its CalcValues theta), the one along `flip`
// reversed (flip < 0: none), then the Cartesian velocity that has them under the metric rows of
// ghost node n: Cramer's rule, terms in the reference's order (src/BC.c:2556-2558, 2697-2699)
__device__ __forceinline__ void ghost_velocity(const BcArgs &a, const long n, const long s, const int flip,
                                               double &u, double &v, double &w)
{
    const double us = a.pr.f[1][s], vs = a.pr.f[2][s], ws = a.pr.f[3][s];
    double t0 = us * a.met[0][s] + vs * a.met[1][s] + ws * a.met[2][s];
    double t1 = us * a.met[3][s] + vs * a.met[4][s] + ws * a.met[5][s];
    double t2 = us * a.met[6][s] + vs * a.met[7][s] + ws * a.met[8][s];
    if (flip == 0) t0 = -t0;
    if (flip == 1) t1 = -t1;
    if (flip == 2) t2 = -t2;
    const double xx = a.met[0][n], xy = a.met[1][n], xz = a.met[2][n], ex = a.met[3][n], ey = a.met[4][n],
                 ez = a.met[5][n], zx = a.met[6][n], zy = a.met[7][n], zz = a.met[8][n];
    const double den = -xx * ey * zz + xx * ez * zy + xy * ex * zz - xy * ez * zx - xz * ex * zy + xz * ey * zx;
    u = -(t0 * ey * zz - t0 * ez * zy - t1 * xy * zz + t1 * xz * zy + t2 * xy * ez - t2 * xz * ey) / den;
    v = -(-t0 * ex * zz + t0 * ez * zx + t1 * xx * zz - t1 * xz * zx - t2 * xx * ez + t2 * xz * ex) / den;
    w = -(-t0 * ex * zy + t0 * ey * zx + t1 * xx * zy - t1 * xy * zx - t2 * xx * ey + t2 * xy * ex) / -den;
}

__device__ __forceinline__ void face_body(const BcArgs &a, const long t)
{
    long n, s;
    if (!face_node(a, t, n, s)) return;
    const Consts &k = a.k;
    if (a.kind == SHX_BC_WALL_SLIP || a.kind == SHX_BC_WALL_NOSLIP || a.kind == SHX_BC_WALL_ISOTHERMAL) {
        // WriteWallSlipBoundary (src/BC.c:2688-2719): the mirror's contravariant velocity along
        // the face's direction reversed; WriteWallNoSlipBoundary (src/BC.c:2657-2686): all
        // velocities reversed.  rho, e, p, c, T and mu from the mirror node.  The corrector these
        // writers store is applied by the WENO sweeps on the ghost nodes of wall faces.
        double u, v, w;
        if (a.kind == SHX_BC_WALL_SLIP) {
            ghost_velocity(a, n, s, a.dir, u, v, w);
        } else {
            u = -a.pr.f[1][s]; v = -a.pr.f[2][s]; w = -a.pr.f[3][s];
        }
        a.pr.f[0][n] = a.pr.f[0][s]; a.pr.f[1][n] = u; a.pr.f[2][n] = v; a.pr.f[3][n] = w;
        a.pr.f[4][n] = a.pr.f[4][s];
        if (a.P) { a.P[n] = a.P[s]; a.C[n] = a.C[s]; }
        // WriteWallNoSlipIsothermalBoundary (src/BC.c:2943-2978): T = T_wall on the ghost
        if (a.T) { a.T[n] = (a.kind == SHX_BC_WALL_ISOTHERMAL) ? a.T_wall : a.T[s]; a.mu[n] = a.mu[s]; }
        return;
    }
    if (a.kind == SHX_BC_OUTFLOW_NORMAL || a.kind == SHX_BC_OUTFLOW_RUDY) {
        double u, v, w;
        ghost_velocity(a, n, s, -1, u, v, w);
        double p, rho;
        if (a.kind == SHX_BC_OUTFLOW_RUDY) {
            p = (a.PL[n] + a.alpha_tau * a.p_out + a.RL[n] * a.C0[s] * (u - a.UL[n]) / k.upsilon) / (1.0 + a.alpha_tau);
            rho = p / a.P[s] * a.pr.f[0][s];
        } else {
            p = a.p_out;
            rho = a.pr.f[0][s];
        }
        a.pr.f[0][n] = rho; a.pr.f[1][n] = u; a.pr.f[2][n] = v; a.pr.f[3][n] = w;
        a.pr.f[4][n] = a.pr.f[4][s];
        a.P[n] = p;
        a.C[n] = a.C[s];
        if (a.T) { a.T[n] = a.T[s]; a.mu[n] = a.mu[s]; }
        return;
    }
    if (a.kind == SHX_BC_FARFIELD) {
        // WriteFarfieldBoundary (src/BC.c:2817-2852): the mirror node's values
#pragma unroll
        for (int m = 0; m < 5; m++) a.pr.f[m][n] = a.pr.f[m][s];
        if (a.P) { a.P[n] = a.P[s]; a.C[n] = a.C[s]; }
        if (a.T) { a.T[n] = a.T[s]; a.mu[n] = a.mu[s]; }
        return;
    }
    if (a.kind == SHX_BC_INFLOW_SUB_NORMAL || a.kind == SHX_BC_INFLOW_SUB_RIEMANN ||
        a.kind == SHX_BC_INFLOW_SUB_ISENTROPIC) {
        // subsonic inflow (src/BC.c:2363-2503): velocity, rho and p by the writer's rule, e and c
        // from them, T and mu from the mirror node
        const double gm = k.gamma;
        double u, v, w, rho, p;
        if (a.kind == SHX_BC_INFLOW_SUB_NORMAL) {
            u = a.u_in; v = a.v_in; w = a.w_in;
            p = a.p_in; rho = a.rho_in;
        } else if (a.kind == SHX_BC_INFLOW_SUB_RIEMANN) {
            // incoming invariant and entropy of the far field, outgoing invariant of the mirror
            const double rs = a.pr.f[0][s], us = a.pr.f[1][s];
            const double cii = a.P ? a.C[s]
                                   : sqrt(k.upsilon * gm * pressure(k, rs, us, a.pr.f[2][s], a.pr.f[3][s], a.pr.f[4][s]) / rs);
            u = 0.5 * (us + a.u_in) + (a.cinf - cii) / (gm - 1.);
            v = a.v_in; w = a.w_in;
            const double d1 = (0.5 * (cii + a.cinf) + 0.25 * (gm - 1.) * (a.u_in - us)) / a.cinf;
            rho = a.rho_in * pow(d1, 2. / (gm - 1.));
            p = a.p_in * pow(d1, 2. * gm / (gm - 1.));
        } else {
            // the mirror's contravariant velocities, rho from Bernoulli along the far field's isentrope
            ghost_velocity(a, n, s, -1, u, v, w);
            rho = pow(pow(a.rho_in, (gm - 1.0)) - pow(a.rho_in, gm) / a.p_in * (gm - 1.0) / gm * gm * a.mach * a.mach
                      * (u * u / 2.0), 1.0 / (gm - 1.0));
            p = a.p_in * pow(rho / a.rho_in, gm);
        }
        a.pr.f[0][n] = rho; a.pr.f[1][n] = u; a.pr.f[2][n] = v; a.pr.f[3][n] = w;
        a.pr.f[4][n] = (0.5 * ((u * u) + (v * v) + (w * w)) + p / rho / (gm - 1.0) * k.upsilon);
        if (a.T) { a.T[n] = a.T[s]; a.mu[n] = a.mu[s]; }
        if (a.P) { a.P[n] = p; a.C[n] = sqrt(k.upsilon * gm * p / rho); }
        return;
    }
    if (a.kind == SHX_BC_MANUFACTURED) {
        // WriteManufacturedSolution: exact state, T = |p/rho|, mu = 1 (c, p and theta are
        // recomputed from rho,u,v,w,e where the flux kernels need them)
        const double rho = a.bs[0][n];
        for (int m = 0; m < 5; m++) a.pr.f[m][n] = a.bs[m][n];
        if (a.T) { a.T[n] = fabs(a.bs[5][n] / rho); a.mu[n] = 1.0; }
        if (a.P) { a.P[n] = a.bs[5][n]; a.C[n] = sqrt(k.upsilon * k.gamma * a.bs[5][n] / rho); }
    } else if (a.kind == SHX_BC_INFLOW_SUPERSONIC) {
        const double rho = a.rho_in, u = a.u_in, v = a.v_in, w = a.w_in, p = a.p_in;
        a.pr.f[0][n] = rho; a.pr.f[1][n] = u; a.pr.f[2][n] = v; a.pr.f[3][n] = w;
        a.pr.f[4][n] = (0.5 * ((u * u) + (v * v) + (w * w)) + p / rho / (k.gamma - 1.0) * k.upsilon);
        if (a.T) {
            a.T[n] = a.p_in / a.rho_in;
            a.mu[n] = ((1.0 + k.suth) * pow(p / rho, 1.5) / (p / rho + k.suth));
        }
        if (a.P) { a.P[n] = p; a.C[n] = sqrt(k.upsilon * k.gamma * p / rho); }
    } else {
        // Riemann invariants: outgoing from the mirror node, incoming from the inflow state
        const double rs = a.pr.f[0][s], us = a.pr.f[1][s], vs = a.pr.f[2][s], ws = a.pr.f[3][s];
        const double ps = a.P ? a.P[s] : pressure(k, rs, us, vs, ws, a.pr.f[4][s]);
        const double cii = a.P ? a.C[s] : sqrt(k.upsilon * k.gamma * ps / rs);
        const double u = 0.5 * (us + a.u_in) - (a.cinf - cii) / (k.gamma - 1.);
        const double d1 = (0.5 * (cii + a.cinf) - 0.25 * (k.gamma - 1.) * (a.u_in - us)) / cii;
        const double rho = rs * pow(d1, 2. / (k.gamma - 1.));
        const double p = ps * pow(d1, 2. * k.gamma / (k.gamma - 1.));
        a.pr.f[0][n] = rho; a.pr.f[1][n] = u; a.pr.f[2][n] = vs; a.pr.f[3][n] = ws;
        a.pr.f[4][n] = (0.5 * ((u * u) + (vs * vs) + (ws * ws)) + p / rho / (k.gamma - 1.0) * k.upsilon);
        if (a.T) { a.T[n] = a.T[s]; a.mu[n] = a.mu[s]; }
        if (a.P) { a.P[n] = p; a.C[n] = sqrt(k.upsilon * k.gamma * p / rho); }
    }
}
// mirrorGhostCellsMetric (src/Functions.c:6849-7014): every ghost node of a wall or farfield face's
// slab (the whole tangential range) takes the metric of its mirror node
__global__ void k_mirror_metric(BcArgs a, Fields10 met)
{
    long n, s;
    if (!face_node(a, (long)blockIdx.x * blockDim.x + threadIdx.x, n, s)) return;
#pragma unroll
    for (int m = 0; m < 10; m++) met.f[m][n] = met.f[m][s];
}
__global__ void k_face_save(BcArgs a)
{
    face_save_body(a, (long)blockIdx.x * blockDim.x + threadIdx.x);
}
__global__ void k_face_bc(BcArgs a)
{
    face_body(a, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Ghost nodes of non-periodic directions carried from one stage buffer to the next: the
// reference keeps one U_RK array whose ghost layers persist between stages, which matters where
// a writer reads a ghost another face wrote in the previous stage.
__device__ __forceinline__ void carry_body(const CFields5 &from, const Fields5 &to, const Geom &gm, const long t)
{
    if (t >= gm.ntot) return;
    int idx[3];
    const long n = padded_node(gm, t, idx);
    bool ghost = false;
#pragma unroll
    for (int d = 0; d < 3; d++)
        if (gm.ng[d] > 1 && !gm.per[d] && (idx[d] < gm.lo[d] || idx[d] >= gm.lo[d] + gm.n[d])) ghost = true;
    if (!ghost) return;
#pragma unroll
    for (int m = 0; m < 5; m++) to.f[m][n] = from.f[m][n];
}
__global__ void k_carry_ghosts(CFields5 from, Fields5 to, Geom gm)
{
    carry_body(from, to, gm, (long)blockIdx.x * blockDim.x + threadIdx.x);
}
// every ghost node (periodic directions included) copied from one buffer to another
__global__ void k_copy_ghosts(CFields5 from, Fields5 to, Geom gm)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= gm.ntot) return;
    int idx[3];
    const long n = padded_node(gm, t, idx);
    bool ghost = false;
#pragma unroll
    for (int d = 0; d < 3; d++)
        if (gm.ng[d] > 1 && (idx[d] < gm.lo[d] || idx[d] >= gm.lo[d] + gm.n[d])) ghost = true;
    if (!ghost) return;
#pragma unroll
    for (int m = 0; m < 5; m++) to.f[m][n] = from.f[m][n];
}

// T = |p/rho| and mu (Sutherland, or 1 under the manufactured solution) of one node, CalcValues
// (src/Functions.c:2462-2476)
__device__ __forceinline__ void node_tmu(const Consts &k, const double p, const double rho, double &T, double &mu)
{
    T = fabs(p / rho);
    mu = k.mms ? 1.0 : ((1.0 + k.suth) * pow(p / rho, 1.5) / (p / rho + k.suth));
}

// T and mu of node n (viscous runs), and its stored p and c (cases with a subsonic outflow face)
__device__ __forceinline__ void node_values(const double rho, const double u, const double v, const double w,
                                            const double e, const gdp T, const gdp mu, const gdp P, const gdp C,
                                            const Consts &k, const long n)
{
    const double p = pressure(k, rho, u, v, w, e);
    if (T) {
        double t, m;
        node_tmu(k, p, rho, t, m);
        T[n] = t;
        mu[n] = m;
    }
    if (P) { P[n] = p; C[n] = sqrt(k.upsilon * k.gamma * p / rho); }
}
// CalcValues on all padded nodes
__device__ __forceinline__ void values_body(const CFields5 &pr, const gdp T, const gdp mu, const gdp P,
                                            const gdp C, const Geom &gm, const Consts &k, const long t)
{
    if (t >= gm.ntot) return;
    int idx[3];
    const long n = padded_node(gm, t, idx);
    node_values(pr.f[0][n], pr.f[1][n], pr.f[2][n], pr.f[3][n], pr.f[4][n], T, mu, P, C, k, n);
}
// Periodic ghost fill of the primitives and CalcValues in one pass: a ghost node copies its
// periodic image and evaluates the values from the copied primitives, which gives the bits of the
// fill followed by CalcValues.  Only ghost nodes are written to pr, and they are never an image,
// so no thread reads a node another thread writes.
__device__ __forceinline__ void ghost_values_body(const Fields5 &pr, const gdp T, const gdp mu, const gdp P,
                                                  const gdp C, const Geom &gm, const Consts &k, const long t)
{
    if (t >= gm.ntot) return;
    int idx[3];
    const long n = padded_node(gm, t, idx);
    bool ghost;
    const long src = periodic_image(gm, idx, ghost);
    double q[5];
#pragma unroll
    for (int m = 0; m < 5; m++) q[m] = pr.f[m][src];
    if (ghost) {
#pragma unroll
        for (int m = 0; m < 5; m++) pr.f[m][n] = q[m];
    }
    node_values(q[0], q[1], q[2], q[3], q[4], T, mu, P, C, k, n);
}
__global__ void k_calc_values(CFields5 pr, gdp T, gdp mu, gdp P, gdp C, Geom gm, Consts k)
{
    values_body(pr, T, mu, P, C, gm, k, (long)blockIdx.x * blockDim.x + threadIdx.x);
}
__global__ void k_ghost_values(Fields5 pr, gdp T, gdp mu, gdp P, gdp C, Geom gm, Consts k)
{
    ghost_values_body(pr, T, mu, P, C, gm, k, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- block placement ------------------------------------------------------------------------
// Blocks are dispatched round-robin over the eight XCDs (block p on XCD p % 8), each with its own
// L2.  xcd_order hands XCD x a contiguous range of tiles in dispatch order, so tiles next to each
// other in the tile order (which share halo nodes) run on one L2 at about the same time.
__device__ __forceinline__ int xcd_order(const int p, const int total)
{
    const int per = total >> 3, rem = total & 7, x = p & 7, s = p >> 3;
    return x * per + min(x, rem) + s;
}
// Tile of block p; tg = (A tiles, B tiles, sweep tiles, order).  order 1: sweep tiles fastest,
// then A, then B, placed by xcd_order (neighbours along the sweep share halo nodes in one L2);
// order 0: A fastest in launch order, which keeps the row segments that consecutive blocks read
// adjacent in memory (measured faster where the sweep axis is the slowest one in memory).
__device__ __forceinline__ void sweep_tile(const int p, const int tg[4], int &bx, int &by, int &bz)
{
    if (tg[3] == 0) {
        bx = p % tg[0];
        by = (p / tg[0]) % tg[1];
        bz = p / (tg[0] * tg[1]);
        return;
    }
    const int q = xcd_order(p, tg[0] * tg[1] * tg[2]);
    bz = q % tg[2];
    const int r = q / tg[2];
    bx = r % tg[0];
    by = r / tg[0];
}

// ---- WENO sweep ---------------------------------------------------------------------------
struct WenoArgs {
    CFields5 pr;          // rho, u, v, w, e of the stage input (ghosts filled)
    gcdp m0, m1, m2, J;               // metric row of the sweep direction and the Jacobian
                                      // (diagonal metrics: m0 is the row's diagonal entry)
    Fields5 q;            // stage residual, interior nodes
    gcdp P, C;            // stored p and c, read on the ghost nodes of subsonic outflow faces
    int pc_lo, pc_hi;     // the lower / upper face of the sweep direction stores p and c on its ghosts
    int w_lo, w_hi;       // the lower / upper face of the sweep direction is a wall
    double c_lo[3], c_hi[3];  // its corrector of the convective momentum terms (src/WENO.c:265-277)
    int tg[4];            // tile grid (A tiles, B tiles, sweep tiles) and order of a single-case launch
    long sD, sA, sB;      // strides: sweep, fast transverse, slow transverse
    int nD, nA, nB;       // interior counts
    int oD, oA, oB;       // padded index of the first interior node per axis
    int ngD;              // padded extent along the sweep
    int first;            // 1: Q = -div (first direction of the stage), 0: Q -= div
    int coalesce_d;       // load mapping: 1 when the sweep axis is contiguous in memory
    int alignA;           // tiles start at padded A index 0 (A axis contiguous, pitch padded)
    Consts k;
};

// per-node values staged in LDS
enum { V_F0, V_F1, V_F2, V_F3, V_F4, V_U0, V_U1, V_U2, V_U3, V_U4, V_LAM, V_U, V_V, V_W, V_C,
       V_NX, V_NY, V_NZ, V_J, NV };

// AX = -1: general metric rows; AX = 0, 1, 2: diagonal metric, the sweep's normal is the axis AX
// (the reference's 3D metric, src/Functions.c:6479-6484), which drops two metric loads, two LDS
// fields and the zero entries of the eigenvector rows.
// MIX: mixed-precision form.  Node fluxes, the central part Theta, the eigen-system and the
// divergence stay FP64; the characteristic projection (of node-to-node differences of F and U,
// formed in FP64 and rounded once), the smoothness indicators, weights and the deviation Phi, and
// R . Phi are FP32.  Phi is a correction on top of Theta, so its rounding enters at the size of
// the correction rather than of the flux.
// LDS of one WENO sweep block, in doubles: the node window (NVK fields) and, in the mixed form,
// room for two of the ten FP32 difference arrays (the other eight reuse four conserved-variable
// fields of the window).  Diagonal metric: no normal components; the sign of the
// metric entry rides on V_LAM (the spectral radius is positive) and J moves to V_NX, which keeps
// the 8 x 32 tile's window under 40 KB (four blocks per CU).
// DIFF (FP64 WENO5): the F and U arrays hold node-to-node differences once the node pass is
// done, and one more array keeps rho for the interface state (below).
template <int S, int AX, bool MIX>
constexpr bool weno_diff() { return S == 5 && !MIX; }
template <int S, int TA, int TDI, int AX, bool MIX>
constexpr int weno_lds()
{
    constexpr int NW = (TDI - 1 + (S + 1)) * TA;
    constexpr int NVK = (AX < 0) ? NV : NV - 3;
    return NVK * NW + ((MIX || weno_diff<S, AX, MIX>()) ? NW : 0);
}

// KEEP: leave Q alone and return this sweep's contribution of the thread's node in keep[5] (a
// fused kernel adds it later); otherwise Q = -div (first sweep of the stage) or Q -= div.
template <int S, int TA, int TDI, int AX, bool MIX, bool KEEP = false>
__device__ __forceinline__ void weno_body(const WenoArgs &a, const int bx, const int by, const int bz,
                                          double *lds, double *keep = nullptr)
{
    constexpr int G = (S + 1) / 2;
    constexpr int R = G - 1;
    constexpr int WD = TDI - 1 + 2 * G;   // window nodes along the sweep
    constexpr int NW = WD * TA;
    constexpr int NT = TA * TDI;
    constexpr int NVK = (AX < 0) ? NV : NV - 3;
    constexpr int VJ = (AX < 0) ? V_J : V_NX;
    // DIFF (FP64 WENO5, any metric): after the node pass, position o of the F and U arrays holds
    // node o+1 minus node o.
    // The projection then runs on the five differences the indicators use (instead of six node
    // values, whose differences were taken after projecting), and the central part Theta is
    // written in differences: Theta = F[oc] + (D[oc-1] + 6 D[oc] - D[oc+1]) / 12, whose F[oc]
    // terms telescope in the divergence (node oc+1's divergence = D[oc] + the difference of the
    // two interface values without F).  rho, for the interface state, keeps its own array VRHO.
    constexpr bool DIFF = weno_diff<S, AX, MIX>();
    constexpr int VRHO = NVK;
    double *const sm = lds;                       // NVK * NW (+ NW)
    // mixed form: the FP32 differences of F (arrays 0..4) and U (5..9).  Once formed, the node
    // values of rho u, rho v, rho w and rho e are not read again, so arrays 0..7 take their place
    // (V_U1..V_U4) and only arrays 8 and 9 extend the window: the 8 x 32 tile keeps four blocks
    // per CU.
    auto dfa = [&](const int j) -> float * {
        return (j < 8) ? (float *)(lds + V_U1 * NW) + j * NW : (float *)(lds + NVK * NW) + (j - 8) * NW;
    };

    const int t = threadIdx.x;
    // first padded A index of the tile (aligned to TA when the A axis is contiguous in memory)
    const int pa0 = a.alignA ? bx * TA : bx * TA + a.oA;
    const int b = by;
    const int d0 = bz * (TDI - 1);               // first node of this tile (interior-relative)
    const long bbase = (long)(b + a.oB) * a.sB;
    const Consts &k = a.k;

    // phase 1: per-node flux, conserved vector, spectral radius and eigenvector inputs.  Sweeps
    // of a direction with a wall face take a second copy of the loop (WALL), so the others keep
    // the plain node flux.
    auto node_pass = [&](auto wall_c) {
    constexpr bool WALL = decltype(wall_c)::value;
    for (int idx = t; idx < NW; idx += NT) {
        int aa, dd;
        if (a.coalesce_d) { aa = idx / WD; dd = idx - aa * WD; }
        else { dd = idx / TA; aa = idx - dd * TA; }
        const int pa = max(a.oA, min(pa0 + aa, a.oA + a.nA - 1));
        int dp = d0 - G + dd + a.oD;
        dp = max(0, min(dp, a.ngD - 1));
        const long n = bbase + (long)pa * a.sA + (long)dp * a.sD;
        const double rho = a.pr.f[0][n], u = a.pr.f[1][n], v = a.pr.f[2][n], w = a.pr.f[3][n],
                     e = a.pr.f[4][n];
        const double J = a.J[n];
        double p = pressure_fast(k, rho, u, v, w, e);
        const double ir = frcp(rho);
        double cs = sqrt(k.upsilon * k.gamma * p * ir);
        if ((a.pc_lo && dp < a.oD) || (a.pc_hi && dp >= a.oD + a.nD)) { p = a.P[n]; cs = a.C[n]; }
        const double pu = p * k.upsilon;
        const int o = aa * WD + dd;
        // velocities of the convective momentum terms: times the wall corrector (+-1, exact) on
        // the ghost nodes of a wall face of this direction, the velocities themselves elsewhere
        double uc[3] = {u, v, w};
        if constexpr (WALL) {
            if ((a.w_lo && dp < a.oD) || (a.w_hi && dp >= a.oD + a.nD)) {
                const bool lo = dp < a.oD;
#pragma unroll
                for (int q = 0; q < 3; q++) uc[q] *= lo ? a.c_lo[q] : a.c_hi[q];
            }
        }
        if constexpr (AX < 0) {
            const double mx = a.m0[n], my = a.m1[n], mz = a.m2[n];
            const double th = u * mx + v * my + w * mz;
            const double rt = rho * th;
            sm[V_F0 * NW + o] = J * rt;
            sm[V_F1 * NW + o] = J * (rt * uc[0] + mx * pu);
            sm[V_F2 * NW + o] = J * (rt * uc[1] + my * pu);
            sm[V_F3 * NW + o] = J * (rt * uc[2] + mz * pu);
            sm[V_F4 * NW + o] = J * (rt * (e + pu * ir));
            const double mn = sqrt(mx * mx + my * my + mz * mz);
            const double imn = frcp(mn);
            sm[V_LAM * NW + o] = fabs(th) + cs * mn;
            sm[V_NX * NW + o] = mx * imn;
            sm[V_NY * NW + o] = my * imn;
            sm[V_NZ * NW + o] = mz * imn;
        } else {
            const double md = a.m0[n];
            const double uu[3] = {u, v, w};
            const double th = uu[AX] * md;
            const double rt = rho * th;
            sm[V_F0 * NW + o] = J * rt;
#pragma unroll
            for (int q = 0; q < 3; q++) sm[(V_F1 + q) * NW + o] = (q == AX) ? J * (rt * uc[q] + md * pu) : J * (rt * uc[q]);
            sm[V_F4 * NW + o] = J * (rt * (e + pu * ir));
            const double lam = fabs(th) + cs * fabs(md);
            sm[V_LAM * NW + o] = (md > 0.0) ? lam : -lam;
        }
        sm[V_U0 * NW + o] = rho;
        if constexpr (DIFF) sm[VRHO * NW + o] = rho;
        sm[V_U1 * NW + o] = rho * u;
        sm[V_U2 * NW + o] = rho * v;
        sm[V_U3 * NW + o] = rho * w;
        sm[V_U4 * NW + o] = rho * e;
        sm[V_U * NW + o] = u;
        sm[V_V * NW + o] = v;
        sm[V_W * NW + o] = w;
        sm[V_C * NW + o] = cs;
        sm[VJ * NW + o] = J;
    }
    };
    if (a.w_lo || a.w_hi) node_pass(std::true_type{});
    else node_pass(std::false_type{});
    __syncthreads();
    if constexpr (MIX) {
        // node-to-node differences of F and U, rounded once to FP32 (position o: node o+1 - node o),
        // all read before any is written over the conserved values
        constexpr int PER = (NW + NT - 1) / NT;
        float dv[PER][10];
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int idx = t + r * NT;
            if (idx < NW && idx % WD != WD - 1) {
#pragma unroll
                for (int q = 0; q < 5; q++) {
                    dv[r][q] = (float)(sm[(V_F0 + q) * NW + idx + 1] - sm[(V_F0 + q) * NW + idx]);
                    dv[r][5 + q] = (float)(sm[(V_U0 + q) * NW + idx + 1] - sm[(V_U0 + q) * NW + idx]);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int idx = t + r * NT;
            if (idx < NW && idx % WD != WD - 1) {
#pragma unroll
                for (int j = 0; j < 10; j++) dfa(j)[idx] = dv[r][j];
            }
        }
        __syncthreads();
    }
    if constexpr (DIFF) {
        // node-to-node differences of F and U, all read before any is written over the node values
        constexpr int PER = (NW + NT - 1) / NT;
        double dv[PER][10];
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int idx = t + r * NT;
            if (idx < NW && idx % WD != WD - 1) {
#pragma unroll
                for (int q = 0; q < 10; q++) dv[r][q] = sm[(V_F0 + q) * NW + idx + 1] - sm[(V_F0 + q) * NW + idx];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int idx = t + r * NT;
            if (idx < NW && idx % WD != WD - 1) {
#pragma unroll
                for (int q = 0; q < 10; q++) sm[(V_F0 + q) * NW + idx] = dv[r][q];
            }
        }
        __syncthreads();
    }

    // phase 2: interface flux between nodes (d0-1+di) and (d0+di)
    const int ta = t / TDI, di = t - ta * TDI;
    const int ob = ta * WD + di;                 // window position of the stencil's first node
    const int oc = ob + R;                       // left node of the interface
    double out[5];
    double dcur[5];                              // DIFF: D[oc] of every field, for the divergence
    {
        const double Jh = (sm[VJ * NW + oc] + sm[VJ * NW + oc + 1]) / 2.0;
        const double u = 0.5 * (sm[V_U * NW + oc] + sm[V_U * NW + oc + 1]);
        const double v = 0.5 * (sm[V_V * NW + oc] + sm[V_V * NW + oc + 1]);
        const double w = 0.5 * (sm[V_W * NW + oc] + sm[V_W * NW + oc + 1]);
        constexpr int VR = DIFF ? VRHO : V_U0;
        const double rho = 0.5 * (sm[VR * NW + oc] + sm[VR * NW + oc + 1]);
        const double cs = 0.5 * (sm[V_C * NW + oc] + sm[V_C * NW + oc + 1]);
        double nx, ny, nz;
        if constexpr (AX < 0) {
            nx = 0.5 * (sm[V_NX * NW + oc] + sm[V_NX * NW + oc + 1]);
            ny = 0.5 * (sm[V_NY * NW + oc] + sm[V_NY * NW + oc + 1]);
            nz = 0.5 * (sm[V_NZ * NW + oc] + sm[V_NZ * NW + oc + 1]);
        } else {
            const double ns = 0.5 * (copysign(1.0, sm[V_LAM * NW + oc]) + copysign(1.0, sm[V_LAM * NW + oc + 1]));
            nx = (AX == 0) ? ns : 0.0;
            ny = (AX == 1) ? ns : 0.0;
            nz = (AX == 2) ? ns : 0.0;
        }
        const double nsig = (AX == 0) ? nx : ((AX == 1) ? ny : nz);
        const Eig e = make_eig(nx, ny, nz, u, v, w, rho, cs, k.gamma);
        double lam = 0.0;
#pragma unroll
        for (int s = 0; s < 2 * G; s++) lam = fmax(lam, (AX < 0) ? sm[V_LAM * NW + ob + s] : fabs(sm[V_LAM * NW + ob + s]));
        const double lj = lam * Jh;

        // Characteristic projection per field c: a_s = L_c . F_s and b_s = L_c . U_s on the 2G
        // stencil nodes, f+_s = (a_s + lam J b_s)/2 on nodes m-r..m+r and f-_s = (a - lam J b)/2
        // on the mirrored nodes m+r+1..m-r+1 (src/WENO.c:334-385).
        double phis[5];
        float phif[5];
        auto project = [&](auto cc) {
            constexpr int c = decltype(cc)::value;
            if constexpr (MIX) {
                double Ld[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                if constexpr (AX < 0) left_row<c>(e, Ld);
                else left_row_ax<c, (AX < 0 ? 0 : AX)>(e, nsig, Ld);
                float Lf[5];
#pragma unroll
                for (int q = 0; q < 5; q++) Lf[q] = (float)Ld[q];
                const float ljf = (float)lj;
                // projected differences da_s = L_c . (F_{s+1} - F_s), db_s likewise for U
                float dp[S - 1], dm[S - 1];
                float da[2 * G - 1], db[2 * G - 1];
#pragma unroll
                for (int s = 0; s < 2 * G - 1; s++) {
                    float x = 0.f, y = 0.f;
#pragma unroll
                    for (int q = 0; q < 5; q++) {
                        if (AX < 0 || lnz(c, AX, q)) {
                            x += Lf[q] * dfa(q)[ob + s];
                            y += Lf[q] * dfa(5 + q)[ob + s];
                        }
                    }
                    da[s] = x;
                    db[s] = y;
                }
                // differences of the split fluxes (times 2, see below): f+ on nodes 0..S-1, f- mirrored
#pragma unroll
                for (int s = 0; s < S - 1; s++) {
                    dp[s] = da[s] + ljf * db[s];
                    dm[s] = da[S - 1 - s] - ljf * db[S - 1 - s];
                }
                if constexpr (S == 5) {
                    phif[c] = 0.5f * (-phi_w5_f(dp[0], dp[1], dp[2], dp[3]) + phi_w5_f(dm[0], dm[1], dm[2], dm[3]));
                } else {
                    phif[c] = 0.5f * (-phi_w9_f(dp) + phi_w9_f(dm));
                }
                return;
            } else {
            // re-read the stencil from LDS for every field instead of keeping 60 values live
            // across the five fields: 164 instead of 228 VGPRs, three waves per SIMD
            asm volatile("" ::: "memory");
            double Lc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            if constexpr (AX < 0) left_row<c>(e, Lc);
            else left_row_ax<c, (AX < 0 ? 0 : AX)>(e, nsig, Lc);
            if constexpr (DIFF) {
                // projected differences da_s = L_c . (F_{s+1} - F_s), db_s likewise for U, on the
                // five node pairs of the stencil; f+ differences on pairs 0..3, f- on 4..1
                double da[5], db[5];
#pragma unroll
                for (int s = 0; s < 5; s++) {
                    double aa = 0.0, bb = 0.0;
#pragma unroll
                    for (int q = 0; q < 5; q++) {
                        if (AX < 0 || lnz(c, AX, q)) {
                            aa += Lc[q] * sm[(V_F0 + q) * NW + ob + s];
                            bb += Lc[q] * sm[(V_U0 + q) * NW + ob + s];
                        }
                    }
                    da[s] = aa;
                    db[s] = bb;
                }
                phis[c] = 0.5 * (-phi_w5_fast(k.eps, da[0] + lj * db[0], da[1] + lj * db[1], da[2] + lj * db[2],
                                              da[3] + lj * db[3])
                                 + phi_w5_fast(k.eps, da[4] - lj * db[4], da[3] - lj * db[3], da[2] - lj * db[2],
                                               da[1] - lj * db[1]));
                return;
            }
            double av[2 * G], bv[2 * G];
#pragma unroll
            for (int s = 0; s < 2 * G; s++) {
                double aa = 0.0, bb = 0.0;
#pragma unroll
                for (int q = 0; q < 5; q++) {
                    if (AX < 0 || lnz(c, AX, q)) {
                        aa += Lc[q] * sm[(V_F0 + q) * NW + ob + s];
                        bb += Lc[q] * sm[(V_U0 + q) * NW + ob + s];
                    }
                }
                av[s] = aa;
                bv[s] = bb;
            }
            // The split fluxes are formed without their factor 1/2, which is applied to Phi
            // instead: Phi is homogeneous of degree one in its inputs (the weights depend on
            // ratios of the indicators, rescaled by exact powers of two), so halving the inputs
            // or the output gives the same bits wherever epsilon is far below the indicators.
            double fp[S], fm[S];
#pragma unroll
            for (int s = 0; s < S; s++) {
                fp[s] = av[s] + lj * bv[s];
                fm[s] = av[2 * G - 1 - s] - lj * bv[2 * G - 1 - s];
            }
            // WENO9 (WENO5 takes the difference form above): the indicators need the node values
            if constexpr (S == 9) {
                double dp_[8], dm_[8];
#pragma unroll
                for (int s = 0; s < 8; s++) { dp_[s] = fp[s + 1] - fp[s]; dm_[s] = fm[s] - fm[s + 1]; }
                phis[c] = 0.5 * (-phi_w9(k.eps, fp, dp_) + phi_w9(k.eps, fm, dm_));
            }
            }
        };
        project(std::integral_constant<int, 0>{});
        project(std::integral_constant<int, 1>{});
        project(std::integral_constant<int, 2>{});
        project(std::integral_constant<int, 3>{});
        project(std::integral_constant<int, 4>{});

        auto assemble = [&](auto mm) {
            constexpr int m = decltype(mm)::value;
            const double *F = &sm[(V_F0 + m) * NW];
            double th;
            if constexpr (DIFF) {
                // Theta without F[oc] (added back in the divergence as D[oc], the F difference)
                th = 1. / 12. * ((F[oc - 1] - F[oc + 1]) + 6. * F[oc]);
                dcur[m] = F[oc];
            } else if constexpr (S == 5) {
                th = 1. / 12. * (-1. * (F[oc - 1] + F[oc + 2]) + 7. * (F[oc] + F[oc + 1]));
            } else {
                th = 1. / 840. * (533. * (F[oc + 1] + F[oc]) - 139. * (F[oc + 2] + F[oc - 1])
                                  + 29. * (F[oc + 3] + F[oc - 2]) - 3. * (F[oc + 4] + F[oc - 3]));
            }
            double Rr[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            if constexpr (MIX) {
                if constexpr (AX < 0) right_row<m>(e, Rr);
                else right_row_ax<m, (AX < 0 ? 0 : AX)>(e, nsig, Rr);
                float acc = 0.f;
#pragma unroll
                for (int cc = 0; cc < 5; cc++)
                    if (AX < 0 || rnz(m, AX, cc)) acc += (float)Rr[cc] * phif[cc];
                out[m] = (double)acc + th;
            } else if constexpr (AX < 0) {
                right_row<m>(e, Rr);
                out[m] = phis[0] * Rr[0] + phis[1] * Rr[1] + phis[2] * Rr[2] + phis[3] * Rr[3] + phis[4] * Rr[4] + th;
            } else {
                right_row_ax<m, (AX < 0 ? 0 : AX)>(e, nsig, Rr);
                double acc = 0.0;
#pragma unroll
                for (int cc = 0; cc < 5; cc++)
                    if (rnz(m, AX, cc)) acc = acc + phis[cc] * Rr[cc];
                out[m] = acc + th;
            }
        };
        assemble(std::integral_constant<int, 0>{});
        assemble(std::integral_constant<int, 1>{});
        assemble(std::integral_constant<int, 2>{});
        assemble(std::integral_constant<int, 3>{});
        assemble(std::integral_constant<int, 4>{});
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 5; m++) sm[m * NT + t] = out[m];
    __syncthreads();

    // phase 3: divergence into Q for the tile's TDI-1 nodes
    if (di < TDI - 1) {
        const int dn = d0 + di;
        const int pa = pa0 + ta;
        if (dn < a.nD && pa >= a.oA && pa < a.oA + a.nA) {
            const long n = bbase + (long)pa * a.sA + (long)(dn + a.oD) * a.sD;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                double dv = sm[m * NT + t + 1] - sm[m * NT + t];
                if constexpr (DIFF) dv = dcur[m] + dv;
                if constexpr (KEEP) keep[m] = -dv;
                else a.q.f[m][n] = (a.first ? 0.0 : a.q.f[m][n]) - dv;
            }
        }
    }
}
// Waves per SIMD asked of the compiler: the diagonal W5 xi and eta sweeps otherwise take 134
// VGPRs (three waves per SIMD); at 128 (126 for xi, no spills) they run four and are measured 7-9 %
// faster.  The other variants are left to the compiler (the zeta sweep measured slower under the
// same bound).
template <int S, int AX, bool MIX>
constexpr int weno_min_waves() { return (S == 5 && (AX == 0 || AX == 1) && !MIX) ? 4 : 1; }

template <int S, int TA, int TDI, int AX, bool MIX>
__global__ __launch_bounds__(TA * TDI, (weno_min_waves<S, AX, MIX>())) void k_weno(WenoArgs a)
{
    __shared__ double lds[weno_lds<S, TA, TDI, AX, MIX>()];
    int bx, by, bz;
    sweep_tile(blockIdx.x, a.tg, bx, by, bz);
    weno_body<S, TA, TDI, AX, MIX>(a, bx, by, bz, lds);
}

// ---- viscous fluxes: face-based LDS sweep ------------------------------------------------
// CalcViscidFluxesIn{Xi,Eta,Zeta}DirectionDirectly (src/ZD.c:47-694) restructured by faces:
// the reference evaluates both half-point fluxes of every node; here each face flux of
// direction D is computed once from an LDS pencil and differenced.  Face values: midpoint
// derivatives of u,v,w,T along D and midpoint interpolations of the nodal cross derivatives;
// coefficients K = (mu fac)_face averages; energy work term u_i (F(i+1/2) - F(i-1/2)) and the
// nodal product term are evaluated per node as in the reference.
// Stage update of one interior node (src/Functions.c:2300-2365) with the manufactured source
// (Q += J src): shared by k_rk_update and the last viscous sweep of a stage, which applies it
// directly instead of storing Q.
struct RkArgs {
    CFields5 un, us, src;
    Fields5 out, qsum;
    double a, b, qf, ff, mf, tau;
    int first_stage, has_src, three;
    // end of step (the last stage of shx_run, shx_run_cfl and shx_simulate): checkNAN on the new
    // state and the CFL maximum over it, fused into the update.  nan_cnt: running count of
    // non-finite interior densities (those nodes get rho = u = v = w (= stored p) = 1e6);
    // cfl_red: the maximum of the node CFL sum, as the bits of a positive double
    int *nan_cnt;
    gdp nanP;
    unsigned long long *cfl_red;
    double gamma, upsilon;
};

// Reductions over a wavefront (all 64 lanes active).
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Maxima of non-negative doubles reduce into RED_SLOTS slots (positive doubles order like their
// bit patterns, so an unsigned 64-bit atomic max combines them); a block uses slot
// blockIdx % RED_SLOTS, and the host takes the maximum of the slots.  One slot for a whole
// launch would serialise every wave's atomic on one cache line.
constexpr int RED_SLOTS = 256;
__device__ __forceinline__ void wave_max_to(unsigned long long *out, double v)
{
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0 && v > 0.0) {
        unsigned long long *slot = out + (blockIdx.x & (RED_SLOTS - 1));
        const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
        if (bits > *slot) atomicMax(slot, bits);     // a stale read only costs an extra atomic
    }
}

// CFL sum of one node, sum_d |theta_d| + c |grad xi_d| (the per-node spectral radius of
// src/WENO.c:1560-1615 with unit index spacing); m: the node's metric rows xi_x .. zeta_z.
// Shared by k_cfl_max and the fused end-of-step check, which therefore give the same bits.
// DIAG (the off-diagonal entries are zero): |grad xi_d| = |m_dd|, which is what the general form
// gives, since u m_dd + v 0 + w 0 = u m_dd and sqrt(m_dd^2 + 0 + 0) = |m_dd| exactly (a correctly
// rounded square root of a rounded square is the operand's magnitude).
template <bool DIAG = false>
__device__ __forceinline__ double node_cfl_sum(const double gamma, const double upsilon, const int dims,
                                               const double rho, const double u, const double v, const double w,
                                               const double e, const double m[9])
{
    const double p = fabs(rho * (gamma - 1) / upsilon * (e - 0.5 * ((u * u) + (v * v) + (w * w))));
    const double cs = sqrt(upsilon * gamma * p / rho);
    double s = 0.0;
    // unrolled over three directions with a guard, so that m and uu stay in registers
    if constexpr (DIAG) {
        const double uu[3] = {u, v, w};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (d < dims) {
                const double a = m[4 * d];
                s += fabs(uu[d] * a) + cs * fabs(a);
            }
        }
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (d < dims) {
                const double a = m[3 * d], b = m[3 * d + 1], z = m[3 * d + 2];
                s += fabs(u * a + v * b + w * z) + cs * sqrt(a * a + b * b + z * z);
            }
        }
    }
    return s;
}

// The node's U_n and Q_sum loaded ahead of the update (rk_preload), so that a fused sweep can
// issue those loads before its window and flux work and their latency overlaps that work.
struct RkPre { double un[5], qs[5]; };
__device__ __forceinline__ void rk_preload(const RkArgs &A, const long n, RkPre &p)
{
#pragma unroll
    for (int m = 0; m < 5; m++) {
        p.un[m] = (m == 3 && !A.three) ? 0.0 : A.un.f[m][n];
        p.qs[m] = A.first_stage ? 0.0 : A.qsum.f[m][n];
    }
}

// m9, cfl: the node's metric rows and where its CFL sum goes (read only when A.cfl_red is set);
// DIAG: the metric is diagonal (node_cfl_sum); pre: U_n and Q_sum loaded by rk_preload
template <bool DIAG = false>
__device__ __forceinline__ void rk_point(const RkArgs &A, const long n, const double J, double qv[5],
                                         const double *m9 = nullptr, double *cfl = nullptr,
                                         const RkPre *pre = nullptr)
{
    if (A.has_src) {
        qv[0] += J * A.src.f[0][n];
        qv[1] += J * A.src.f[1][n];
        qv[2] += J * A.src.f[2][n];
        if (A.three) qv[3] += J * A.src.f[3][n];
        qv[4] += J * A.src.f[4][n];
    }
    double qs[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        const double old = A.first_stage ? 0.0 : (pre ? pre->qs[m] : A.qsum.f[m][n]);
        qs[m] = old + A.mf * qv[m];
        if (A.mf != 0.0) A.qsum.f[m][n] = qs[m];
    }
    const double tj = A.tau / J;
    const double rho_n = pre ? pre->un[0] : A.un.f[0][n];
    // b = 0 on every RK4 stage and the first SSP-RK3 stage: U_s is not read there
    const double rho_s = (A.b != 0.0) ? A.us.f[0][n] : 0.0;
    const double rho = A.a * rho_n + A.b * rho_s + tj * (A.qf * qv[0] + A.ff * qs[0]);
    A.out.f[0][n] = rho;
    double nv[5];
    nv[0] = rho;
#pragma unroll
    for (int m = 1; m < 5; m++) {
        if (m == 3 && !A.three) { A.out.f[3][n] = 0.0; nv[3] = 0.0; continue; }
        const double us = (A.b != 0.0) ? A.us.f[m][n] : 0.0;
        nv[m] = (A.a * (pre ? pre->un[m] : A.un.f[m][n]) * rho_n + A.b * us * rho_s
                 + tj * (A.qf * qv[m] + A.ff * qs[m])) / rho;
        A.out.f[m][n] = nv[m];
    }
    if (A.nan_cnt && !isfinite(rho)) {
        // checkNAN (src/Functions.c:5419-5467) on the step's U_RK
        A.out.f[0][n] = 1000000; A.out.f[1][n] = 1000000; A.out.f[2][n] = 1000000; A.out.f[3][n] = 1000000;
        if (A.nanP) A.nanP[n] = 1000000;
        atomicAdd(A.nan_cnt, 1);
    } else if (A.cfl_red && cfl) {
        *cfl = node_cfl_sum<DIAG>(A.gamma, A.upsilon, A.three ? 3 : 2, nv[0], nv[1], nv[2], nv[3], nv[4], m9);
    }
}

struct ViscArgs {
    gcdp u, v, w, T, mu;
    gcdp cr[8];                   // cross derivatives (layout depends on DIAG, see host)
    CFields10 met;
    Fields5 q;
    long sD, sA, sB;
    int nD, nA, nB, oD, oA, oB, ngD;
    int ngA;                      // padded extent of the A axis
    int coalesce_d;
    int alignA;
    int three;
    int bug;                      // zeta stencil k offset taken from j (DIR = 2 only)
    int fuse_rk;                  // last sweep of the stage: apply the stage update (rk)
    int iso_hi;                   // the upper face of the sweep direction is an isothermal wall:
    double gam_iso;               // the last interior node's heat-conduction factor (the mean of
                                  // its own and the wall ghost's, src/ZD.c:193-194)
    int tg[4];                    // tile grid and order of a single-case launch
    RkArgs rk;
    Consts k;
};

// fac_{i,k,b} (CreateViscidMetric, src/Functions.c:6245-6386) from row r (= b) and row a (= D).
template <int I, int K>
__device__ __forceinline__ double fac_ikb(const double J, const double r0, const double r1, const double r2,
                                          const double a0, const double a1, const double a2)
{
    if constexpr (I == 0 && K == 0) return J * (4.0 / 3.0 * r0 * a0 + r1 * a1 + r2 * a2);
    if constexpr (I == 0 && K == 1) return J * (-2.0 / 3.0 * r1 * a0 + r0 * a1);
    if constexpr (I == 0 && K == 2) return J * (-2.0 / 3.0 * r2 * a0 + r0 * a2);
    if constexpr (I == 1 && K == 0) return J * (r1 * a0 - 2.0 / 3.0 * r0 * a1);
    if constexpr (I == 1 && K == 1) return J * (r0 * a0 + 4.0 / 3.0 * r1 * a1 + r2 * a2);
    if constexpr (I == 1 && K == 2) return J * (-2.0 / 3.0 * r2 * a1 + r1 * a2);
    if constexpr (I == 2 && K == 0) return J * (r2 * a0 - 2.0 / 3.0 * r0 * a2);
    if constexpr (I == 2 && K == 1) return J * (r2 * a1 - 2.0 / 3.0 * r1 * a2);
    if constexpr (I == 2 && K == 2) return J * (r0 * a0 + r1 * a1 + 4.0 / 3.0 * r2 * a2);
    return J * (r0 * a0 + r1 * a1 + r2 * a2);   // heat flux (I == 3)
}

template <int DIR> struct Others { static constexpr int b1 = (DIR == 0) ? 1 : 0; static constexpr int b2 = (DIR == 2) ? 1 : 2; };

// face flux (momentum x,y,z and heat) from slot values d[12] (slot k*3+b, T slots 9+b) and the
// metric/mu of the face's two nodes (generic metric)
template <int DIR>
__device__ __forceinline__ void face_flux_generic(const double d[12], const double m0[10], const double m1[10],
                                                  const double mu0, const double mu1, double F[4])
{
    const double J0 = m0[9], J1 = m1[9];
    const double a00 = m0[3 * DIR], a01 = m0[3 * DIR + 1], a02 = m0[3 * DIR + 2];
    const double a10 = m1[3 * DIR], a11 = m1[3 * DIR + 1], a12 = m1[3 * DIR + 2];
#pragma unroll
    for (int i = 0; i < 3; i++) F[i] = 0.0;
    F[3] = 0.0;
#define SHX_FACE_TERM(I, K, B)                                                                    \
    {                                                                                             \
        const double f0 = fac_ikb<I, K>(J0, m0[3 * B], m0[3 * B + 1], m0[3 * B + 2], a00, a01, a02); \
        const double f1 = fac_ikb<I, K>(J1, m1[3 * B], m1[3 * B + 1], m1[3 * B + 2], a10, a11, a12); \
        F[I] += d[K * 3 + B] * (0.5 * (mu0 * f0 + mu1 * f1));                                     \
    }
#define SHX_FACE_ROW(I) \
    SHX_FACE_TERM(I, 0, 0) SHX_FACE_TERM(I, 0, 1) SHX_FACE_TERM(I, 0, 2) \
    SHX_FACE_TERM(I, 1, 0) SHX_FACE_TERM(I, 1, 1) SHX_FACE_TERM(I, 1, 2) \
    SHX_FACE_TERM(I, 2, 0) SHX_FACE_TERM(I, 2, 1) SHX_FACE_TERM(I, 2, 2)
    SHX_FACE_ROW(0)
    SHX_FACE_ROW(1)
    SHX_FACE_ROW(2)
#undef SHX_FACE_ROW
#undef SHX_FACE_TERM
#pragma unroll
    for (int b = 0; b < 3; b++) {
        const double f0 = fac_ikb<3, 0>(J0, m0[3 * b], m0[3 * b + 1], m0[3 * b + 2], a00, a01, a02);
        const double f1 = fac_ikb<3, 0>(J1, m1[3 * b], m1[3 * b + 1], m1[3 * b + 2], a10, a11, a12);
        F[3] += d[9 + b] * (0.5 * (mu0 * f0 + mu1 * f1));
    }
}

// diagonal metric (the reference's 3D metric): node values mD (row D, component D), mb[3]
// (diagonal entries of all rows), J.  Only the non-zero factors of the generic table remain.
template <int DIR>
__device__ __forceinline__ void face_flux_diag(const double d[12], const double g0[3], const double J0,
                                               const double g1[3], const double J1, const double mu0,
                                               const double mu1, double F[4])
{
    const double mD0 = g0[DIR], mD1 = g1[DIR];
    const double KA = 0.5 * (mu0 * (J0 * (mD0 * mD0)) + mu1 * (J1 * (mD1 * mD1)));
    const double K43 = 0.5 * (mu0 * (J0 * (4.0 / 3.0 * mD0 * mD0)) + mu1 * (J1 * (4.0 / 3.0 * mD1 * mD1)));
    double FD = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (k == DIR) { FD += d[k * 3 + k] * K43; continue; }
        const double KC = 0.5 * (mu0 * (J0 * (-2.0 / 3.0 * g0[k] * mD0)) + mu1 * (J1 * (-2.0 / 3.0 * g1[k] * mD1)));
        FD += d[k * 3 + k] * KC;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i == DIR) { F[i] = FD; continue; }
        const double KB = 0.5 * (mu0 * (J0 * (g0[i] * mD0)) + mu1 * (J1 * (g1[i] * mD1)));
        F[i] = d[DIR * 3 + i] * KB + d[i * 3 + DIR] * KA;
    }
    F[3] = d[9 + DIR] * KA;
}

// nodal product term sum_blk g_{blk,D} sum_s g_s mu fac_{blk,s} (src/ZD.c:230-266)
template <int DIR>
__device__ __forceinline__ double node_prod_generic(const double g[9], const double m[10], const double mu)
{
    const double J = m[9], a0 = m[3 * DIR], a1 = m[3 * DIR + 1], a2 = m[3 * DIR + 2];
    double prod = 0.0;
#define SHX_PROD_TERM(I, K, B) \
    prod += g[I * 3 + DIR] * g[K * 3 + B] * mu * fac_ikb<I, K>(J, m[3 * B], m[3 * B + 1], m[3 * B + 2], a0, a1, a2);
#define SHX_PROD_ROW(I) \
    SHX_PROD_TERM(I, 0, 0) SHX_PROD_TERM(I, 0, 1) SHX_PROD_TERM(I, 0, 2) \
    SHX_PROD_TERM(I, 1, 0) SHX_PROD_TERM(I, 1, 1) SHX_PROD_TERM(I, 1, 2) \
    SHX_PROD_TERM(I, 2, 0) SHX_PROD_TERM(I, 2, 1) SHX_PROD_TERM(I, 2, 2)
    SHX_PROD_ROW(0)
    SHX_PROD_ROW(1)
    SHX_PROD_ROW(2)
#undef SHX_PROD_ROW
#undef SHX_PROD_TERM
    return prod;
}

template <int DIR>
__device__ __forceinline__ double node_prod_diag(const double g[9], const double gm[3], const double J, const double mu)
{
    const double mD = gm[DIR];
    double prod = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double gd = g[i * 3 + DIR];
        if (i == DIR) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double f = (k == DIR) ? J * (4.0 / 3.0 * mD * mD) : J * (-2.0 / 3.0 * gm[k] * mD);
                prod += gd * g[k * 3 + k] * mu * f;
            }
        } else {
            prod += gd * g[DIR * 3 + i] * mu * (J * (gm[i] * mD));
            prod += gd * g[i * 3 + DIR] * mu * (J * (mD * mD));
        }
    }
    return prod;
}

// LDS slots
enum { VS_U, VS_V, VS_W, VS_T, VS_MU, VS_CR };

// LDS of one viscous sweep block, in doubles: the node window (NVS fields), the face hand-over
// buffer where a tile row straddles wavefronts, and the zeta-index quirk's line values.  Plus-side
// face values go from each thread to the previous one (the minus side stays in registers); where a
// tile row never straddles two wavefronts (TDI divides or is a multiple of the 64-lane wavefront)
// the hand-over is a lane shuffle.
template <int S, int DIR, int TA, int TDI, bool DIAG>
constexpr int visc_lds()
{
    constexpr int G = (S + 1) / 2, NW = (TDI - 1 + 2 * G) * TA, NT = TA * TDI;
    constexpr int NC = DIAG ? 4 : 8, NM = DIAG ? 4 : 10;
    constexpr int NLV = 4 + NC, NP = 2 * G + 1;
    constexpr bool SHFL = (64 % TDI == 0) || (TDI % 64 == 0);
    return (VS_CR + NC + NM) * NW + (SHFL ? 0 : 4 * NT) + ((DIR == 2) ? TA * NP * NLV + TA * 2 * NLV : 0);
}

// ADD: add add[5] (a fused WENO sweep's contribution of the thread's node) to Q before the
// viscous terms.  PRE: load the node's Q (and U_n, Q_sum for the stage update) before the window.
template <int S, int DIR, int TA, int TDI, bool DIAG, bool ADD = false, bool PRE = false>
__device__ __forceinline__ void visc_body(const ViscArgs &a, const int bx, const int by, const int bz,
                                          double *lds, const double *add = nullptr)
{
    constexpr int G = (S + 1) / 2;
    constexpr int WD = TDI - 1 + 2 * G;
    constexpr int NW = WD * TA;
    constexpr int NT = TA * TDI;
    constexpr int NC = DIAG ? 4 : 8;                 // cross-derivative fields
    constexpr int NM = DIAG ? 4 : 10;                // metric values per node
    constexpr int VS_M = VS_CR + NC;
    constexpr int NVS = VS_M + NM;
    constexpr int b1 = Others<DIR>::b1, b2 = Others<DIR>::b2;
    constexpr int NP = 2 * G + 1;                    // line positions used by the zeta-index quirk
    constexpr int NLV = 4 + NC;                      // values per position: u,v,w,T + cross
    constexpr bool SHFL = (64 % TDI == 0) || (TDI % 64 == 0);
    double *const sm = lds;
    double *const fo = sm + NVS * NW;
    double *const line = fo + (SHFL ? 0 : 4 * NT);
    double *const lined = line + TA * NP * NLV;

    const int t = threadIdx.x;
    // first padded A index of the tile (aligned to TA when the A axis is contiguous in memory)
    const int pa0 = a.alignA ? bx * TA : bx * TA + a.oA;
    const int b = by;
    const int d0 = bz * (TDI - 1);
    const long bbase = (long)(b + a.oB) * a.sB;
    const Consts &k = a.k;
    const int ta = t / TDI, di = t - ta * TDI;
    const int dn = d0 + di, pa = pa0 + ta;
    const bool own = di < TDI - 1 && dn < a.nD && pa >= a.oA && pa < a.oA + a.nA;
    const long n = bbase + (long)pa * a.sA + (long)(dn + a.oD) * a.sD;
    // PRE: the node's Q, and U_n and Q_sum where the sweep applies the stage update, are loaded
    // before the window, so that their latency overlaps the staging and the face fluxes
    double q0[5];
    RkPre pre;
    if (PRE && own) {
#pragma unroll
        for (int m = 0; m < 5; m++) q0[m] = a.q.f[m][n];
        if (a.fuse_rk) rk_preload(a.rk, n, pre);
    }

    for (int idx = t; idx < NW; idx += NT) {
        int aa, dd;
        if (a.coalesce_d) { aa = idx / WD; dd = idx - aa * WD; }
        else { dd = idx / TA; aa = idx - dd * TA; }
        const int pa = max(a.oA, min(pa0 + aa, a.oA + a.nA - 1));
        int dp = d0 - G + dd + a.oD;
        dp = max(0, min(dp, a.ngD - 1));
        const long n = bbase + (long)pa * a.sA + (long)dp * a.sD;
        const int o = aa * WD + dd;
        sm[VS_U * NW + o] = a.u[n];
        sm[VS_V * NW + o] = a.v[n];
        sm[VS_W * NW + o] = a.w[n];
        sm[VS_T * NW + o] = a.T[n];
        sm[VS_MU * NW + o] = a.mu[n];
#pragma unroll
        for (int c = 0; c < NC; c++) sm[(VS_CR + c) * NW + o] = a.cr[c][n];
        if constexpr (DIAG) {
            sm[(VS_M + 0) * NW + o] = a.met.f[0][n];
            sm[(VS_M + 1) * NW + o] = a.met.f[4][n];
            sm[(VS_M + 2) * NW + o] = a.met.f[8][n];
            sm[(VS_M + 3) * NW + o] = a.met.f[9][n];
        } else {
#pragma unroll
            for (int c = 0; c < 10; c++) sm[(VS_M + c) * NW + o] = a.met.f[c][n];
        }
    }
    const bool bug = (DIR == 2) && a.bug;
    if (bug) {
        // the reference's zeta stencil (src/ZD.c:547-548) reads positions k' = j + l - g on the
        // (i, j) line for every k: the face values are constant along the line
        for (int idx = t; idx < TA * NP * NLV; idx += NT) {
            const int ln = idx / (NP * NLV);
            const int rem = idx - ln * NP * NLV;
            const int pos = rem / NLV, f = rem - pos * NLV;
            const int jpad = max(a.oA, min(pa0 + ln, a.oA + a.nA - 1));   // A axis is j for DIR = 2
            // the reference's index (i ng_j + j) ng_k + k' with k' = j - g + pos in its unpadded
            // k-fastest layout: where k' passes the row end (ng_j > ng_k) it reads the next rows
            // (or the next plane's first ones), which the padded layout finds at (j', k'') below
            const int kq = jpad - G + pos;
            long n;
            if (kq < a.ngD) {
                n = bbase + (long)jpad * a.sA + kq;
            } else {
                int jr = jpad + kq / a.ngD;
                const int kr = kq % a.ngD;
                const int ia = jr / a.ngA;
                jr -= ia * a.ngA;
                n = bbase + (long)ia * a.sB + (long)jr * a.sA + kr;
            }
            double val;
            if (f == 0) val = a.u[n];
            else if (f == 1) val = a.v[n];
            else if (f == 2) val = a.w[n];
            else if (f == 3) val = a.T[n];
            else val = a.cr[f - 4][n];
            line[(ln * NP + pos) * NLV + f] = val;
        }
    }
    __syncthreads();
    if (bug && di < 2) {
        // di = 0: minus-side values (positions 0..S), di = 1: plus side (positions 1..S+1)
        const double *Lp = &line[(ta * NP + di) * NLV];
#pragma unroll
        for (int f = 0; f < NLV; f++) {
            double acc = 0.0;
            const double *cf = (f < 4) ? k.zd_mid : k.zd_int;
#pragma unroll
            for (int l = 0; l <= S; l++) acc += cf[l] * Lp[l * NLV + f];
            lined[(ta * 2 + di) * NLV + f] = acc;
        }
    }
    if (bug) __syncthreads();

    const int ob = ta * WD + di;
    const int oc = ob + G - 1;
    double out[8];
    {
        const double mu0 = sm[VS_MU * NW + oc], mu1 = sm[VS_MU * NW + oc + 1];
        // slot layout d[k*3+b], T at 9+b
        auto face = [&](const double *own4, const double *cr, double F[4]) {
            double d[12];
#pragma unroll
            for (int s = 0; s < 12; s++) d[s] = 0.0;
            d[0 * 3 + DIR] = own4[0]; d[1 * 3 + DIR] = own4[1]; d[2 * 3 + DIR] = own4[2]; d[9 + DIR] = own4[3];
            if constexpr (DIAG) {
                d[b1 * 3 + b1] = cr[0]; d[DIR * 3 + b1] = cr[1];
                d[b2 * 3 + b2] = cr[2]; d[DIR * 3 + b2] = cr[3];
                double g0[3], g1[3];
                g0[0] = sm[(VS_M + 0) * NW + oc]; g0[1] = sm[(VS_M + 1) * NW + oc]; g0[2] = sm[(VS_M + 2) * NW + oc];
                g1[0] = sm[(VS_M + 0) * NW + oc + 1]; g1[1] = sm[(VS_M + 1) * NW + oc + 1]; g1[2] = sm[(VS_M + 2) * NW + oc + 1];
                face_flux_diag<DIR>(d, g0, sm[(VS_M + 3) * NW + oc], g1, sm[(VS_M + 3) * NW + oc + 1], mu0, mu1, F);
            } else {
                d[0 * 3 + b1] = cr[0]; d[1 * 3 + b1] = cr[1]; d[2 * 3 + b1] = cr[2]; d[9 + b1] = cr[3];
                d[0 * 3 + b2] = cr[4]; d[1 * 3 + b2] = cr[5]; d[2 * 3 + b2] = cr[6]; d[9 + b2] = cr[7];
                double m0[10], m1[10];
#pragma unroll
                for (int c = 0; c < 10; c++) { m0[c] = sm[(VS_M + c) * NW + oc]; m1[c] = sm[(VS_M + c) * NW + oc + 1]; }
                face_flux_generic<DIR>(d, m0, m1, mu0, mu1, F);
            }
        };
        if (!bug) {
            double own4[4] = {0.0, 0.0, 0.0, 0.0}, cr[NC];
#pragma unroll
            for (int c = 0; c < NC; c++) cr[c] = 0.0;
#pragma unroll
            for (int l = 0; l <= S; l++) {
                const double cm = k.zd_mid[l], ci = k.zd_int[l];
                own4[0] += cm * sm[VS_U * NW + ob + l];
                own4[1] += cm * sm[VS_V * NW + ob + l];
                own4[2] += cm * sm[VS_W * NW + ob + l];
                own4[3] += cm * sm[VS_T * NW + ob + l];
#pragma unroll
                for (int c = 0; c < NC; c++) cr[c] += ci * sm[(VS_CR + c) * NW + ob + l];
            }
            double F[4];
            face(own4, cr, F);
#pragma unroll
            for (int m = 0; m < 4; m++) { out[m] = F[m]; out[4 + m] = F[m]; }
        } else {
            double FA[4], FB[4];
            face(&lined[(ta * 2 + 0) * NLV], &lined[(ta * 2 + 0) * NLV + 4], FA);
            face(&lined[(ta * 2 + 1) * NLV], &lined[(ta * 2 + 1) * NLV + 4], FB);
#pragma unroll
            for (int m = 0; m < 4; m++) { out[m] = FA[m]; out[4 + m] = FB[m]; }
        }
    }
    double nxt[4];
    if constexpr (SHFL) {
#pragma unroll
        for (int m = 0; m < 4; m++) nxt[m] = __shfl_down(out[4 + m], 1);
    } else {
#pragma unroll
        for (int m = 0; m < 4; m++) fo[m * NT + t] = out[4 + m];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 4; m++) nxt[m] = (di < TDI - 1) ? fo[m * NT + t + 1] : 0.0;
    }

    double cflv = 0.0;                           // CFL sum of this thread's node (end-of-step check)
    if (di < TDI - 1) {
        if (own) {
            const int p = ob + G;                    // window position of this node
            // faces: minus = this thread's face (A values), plus = next thread's face (B values)
            double dF[4];
#pragma unroll
            for (int m = 0; m < 4; m++) dF[m] = nxt[m] - out[m];
            // nodal gradients: own-direction derivatives from the window, cross ones stored
            double g[9];
#pragma unroll
            for (int s = 0; s < 9; s++) g[s] = 0.0;
            {
                double du = 0.0, dv = 0.0, dw = 0.0;
#pragma unroll
                for (int l = 0; l <= S + 1; l++) {
                    du += k.zd_der[l] * sm[VS_U * NW + ob + l];
                    dv += k.zd_der[l] * sm[VS_V * NW + ob + l];
                    dw += k.zd_der[l] * sm[VS_W * NW + ob + l];
                }
                g[0 * 3 + DIR] = du; g[1 * 3 + DIR] = dv; g[2 * 3 + DIR] = a.three ? dw : 0.0;
            }
            const double mu = sm[VS_MU * NW + p];
            double prod;
            if constexpr (DIAG) {
                g[b1 * 3 + b1] = sm[(VS_CR + 0) * NW + p]; g[DIR * 3 + b1] = sm[(VS_CR + 1) * NW + p];
                g[b2 * 3 + b2] = sm[(VS_CR + 2) * NW + p]; g[DIR * 3 + b2] = sm[(VS_CR + 3) * NW + p];
                const double gmv[3] = {sm[(VS_M + 0) * NW + p], sm[(VS_M + 1) * NW + p], sm[(VS_M + 2) * NW + p]};
                prod = node_prod_diag<DIR>(g, gmv, sm[(VS_M + 3) * NW + p], mu);
            } else {
                g[0 * 3 + b1] = sm[(VS_CR + 0) * NW + p]; g[1 * 3 + b1] = sm[(VS_CR + 1) * NW + p];
                g[2 * 3 + b1] = sm[(VS_CR + 2) * NW + p];
                g[0 * 3 + b2] = sm[(VS_CR + 4) * NW + p]; g[1 * 3 + b2] = sm[(VS_CR + 5) * NW + p];
                g[2 * 3 + b2] = sm[(VS_CR + 6) * NW + p];
                double m[10];
#pragma unroll
                for (int c = 0; c < 10; c++) m[c] = sm[(VS_M + c) * NW + p];
                prod = node_prod_generic<DIR>(g, m, mu);
            }
            const double work = sm[VS_U * NW + p] * dF[0] + sm[VS_V * NW + p] * dF[1] + sm[VS_W * NW + p] * dF[2];
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const double qn = PRE ? q0[m] : a.q.f[m][n];
                q0[m] = ADD ? qn + add[m] : qn;
            }
            if (a.fuse_rk) {
                double qv[5];
                qv[0] = q0[0];
                qv[1] = q0[1] + k.psi * dF[0];
                qv[2] = q0[2] + k.psi * dF[1];
                qv[3] = a.three ? q0[3] + k.psi * dF[2] : q0[3];
                const double gh = (a.iso_hi && dn == a.nD - 1) ? a.gam_iso : k.gam_heat;
                qv[4] = q0[4] + (gh * dF[3] + k.psi * (work + prod));
                const double J = DIAG ? sm[(VS_M + 3) * NW + p] : sm[(VS_M + 9) * NW + p];
                if (a.rk.cfl_red) {
                    // end of step: the node's metric rows for its CFL sum
                    double m9[9];
                    if constexpr (DIAG) {
#pragma unroll
                        for (int c = 0; c < 9; c++) m9[c] = 0.0;
                        m9[0] = sm[(VS_M + 0) * NW + p]; m9[4] = sm[(VS_M + 1) * NW + p]; m9[8] = sm[(VS_M + 2) * NW + p];
                    } else {
#pragma unroll
                        for (int c = 0; c < 9; c++) m9[c] = sm[(VS_M + c) * NW + p];
                    }
                    rk_point<DIAG>(a.rk, n, J, qv, m9, &cflv, PRE ? &pre : nullptr);
                } else {
                    rk_point(a.rk, n, J, qv, nullptr, nullptr, PRE ? &pre : nullptr);
                }
            } else {
                if (ADD) a.q.f[0][n] = q0[0];
                a.q.f[1][n] = q0[1] + k.psi * dF[0];
                a.q.f[2][n] = q0[2] + k.psi * dF[1];
                if (a.three) a.q.f[3][n] = q0[3] + k.psi * dF[2];
                else if (ADD) a.q.f[3][n] = q0[3];
                const double gh = (a.iso_hi && dn == a.nD - 1) ? a.gam_iso : k.gam_heat;
                a.q.f[4][n] = q0[4] + (gh * dF[3] + k.psi * (work + prod));
            }
        }
    }
    if (a.rk.cfl_red) wave_max_to(a.rk.cfl_red, cflv);
}
template <int S, int DIR, int TA, int TDI, bool DIAG>
__global__ __launch_bounds__(TA * TDI) void k_visc_sweep(ViscArgs a)
{
    __shared__ double lds[visc_lds<S, DIR, TA, TDI, DIAG>()];
    int bx, by, bz;
    sweep_tile(blockIdx.x, a.tg, bx, by, bz);
    visc_body<S, DIR, TA, TDI, DIAG>(a, bx, by, bz, lds);
}

// ---- fused zeta direction: WENO and viscous sweeps and the stage update ----------------------
// One block sweeps a 4 x 64 tile along zeta twice.  The WENO phase keeps each node's inviscid
// divergence in registers; after a barrier the viscous phase reloads its window into the same LDS,
// adds that divergence and its own terms to Q and applies the stage update (or stores Q).  Q of
// this direction is read once instead of read-written by two kernels, and blocks in their WENO
// phase (VALU-bound) overlap blocks in their viscous phase (memory-bound).  Used for 3D cases with
// the diagonal metric whose zeta WENO tile is 4 x 64 (the viscous zeta tile), viscous runs only.
struct ZetaArgs { WenoArgs w; ViscArgs v; };
template <int S, bool MIX>
constexpr int zeta_lds()
{
    return weno_lds<S, 4, 64, 2, MIX>() > visc_lds<S, 2, 4, 64, true>() ? weno_lds<S, 4, 64, 2, MIX>()
                                                                       : visc_lds<S, 2, 4, 64, true>();
}
template <int S, bool MIX>
__device__ __forceinline__ void zeta_body(const ZetaArgs &z, const int bx, const int by, const int bz, double *lds)
{
    double inv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    weno_body<S, 4, 64, 2, MIX, true>(z.w, bx, by, bz, lds, inv);
    __syncthreads();
    // Q, U_n and Q_sum loaded ahead (measured at 256^3: 1.95 -> 1.75 ms; the same in the fused eta
    // kernel, which stores Q instead of updating the stage, made it slower)
    visc_body<S, 2, 4, 64, true, true, true>(z.v, bx, by, bz, lds, inv);
}
// FP64 W5: held to 128 VGPRs (four waves per SIMD, as the LDS allows); measured 11 % faster than
// the compiler's own 136
template <int S, bool MIX>
__global__ __launch_bounds__(256, ((S == 5 && !MIX) ? 4 : 1)) void k_zeta(ZetaArgs z)
{
    __shared__ double lds[zeta_lds<S, MIX>()];
    int bx, by, bz;
    sweep_tile(blockIdx.x, z.w.tg, bx, by, bz);
    zeta_body<S, MIX>(z, bx, by, bz, lds);
}

// ---- fused eta direction: WENO and viscous eta sweeps in one kernel ----------------------------
// One block sweeps an 8 x 32 tile along eta twice, as the fused zeta kernel does along zeta: the
// WENO phase keeps each node's inviscid divergence in registers, the viscous phase (after a barrier,
// in the same LDS) adds it and its own terms to Q, which this direction reads and writes once.
// 3D diagonal-metric viscous WENO5 cases, where the viscous xi terms ride on the derivative march.
struct EtaArgs { WenoArgs w; ViscArgs v; };
template <bool MIX>
constexpr int eta_lds()
{
    return weno_lds<5, 8, 32, 1, MIX>() > visc_lds<5, 1, 8, 32, true>() ? weno_lds<5, 8, 32, 1, MIX>()
                                                                       : visc_lds<5, 1, 8, 32, true>();
}
template <bool MIX>
__device__ __forceinline__ void eta_body(const EtaArgs &z, const int bx, const int by, const int bz, double *lds)
{
    double inv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    weno_body<5, 8, 32, 1, MIX, true>(z.w, bx, by, bz, lds, inv);
    __syncthreads();
    visc_body<5, 1, 8, 32, true, true>(z.v, bx, by, bz, lds, inv);
}
// FP64: asked for three waves per SIMD.  With the WENO phase in difference form it takes 124
// VGPRs, so four fit (1.44 ms at 256^3; 1.53-1.57 ms at 140 VGPRs and three waves); forcing four
// when it needed 140 spilled and measured 1.82 ms.
constexpr int ETA_WAVES = 3;
template <bool MIX>
__global__ __launch_bounds__(256, (MIX ? 1 : ETA_WAVES)) void k_eta(EtaArgs z)
{
    __shared__ double lds[eta_lds<MIX>()];
    int bx, by, bz;
    sweep_tile(blockIdx.x, z.w.tg, bx, by, bz);
    eta_body<MIX>(z, bx, by, bz, lds);
}

// Stage update (src/Functions.c:2300-2365) + manufactured source: pointwise.
struct UpdArgs {
    Fields5 q;
    gcdp J;
    CFields10 met;        // metric rows, read for the CFL sum of the end-of-step check
    Geom gm;
    RkArgs rk;
};

__device__ __forceinline__ void rk_body(const UpdArgs &A, const long t)
{
    const Geom &gm = A.gm;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    double cfl = 0.0;
    if (t < nint) {
        const int kk = (int)(t % gm.n[2]);
        const long r = t / gm.n[2];
        const int jj = (int)(r % gm.n[1]);
        const int ii = (int)(r / gm.n[1]);
        const long n = (long)(ii + gm.lo[0]) * gm.st[0] + (long)(jj + gm.lo[1]) * gm.st[1] + kk + gm.lo[2];
        double qv[5];
#pragma unroll
        for (int m = 0; m < 5; m++) qv[m] = A.q.f[m][n];
        double m9[9];
#pragma unroll
        for (int c = 0; c < 9; c++) m9[c] = A.rk.cfl_red ? A.met.f[c][n] : 0.0;
        rk_point(A.rk, n, A.J[n], qv, m9, &cfl);
    }
    if (A.rk.cfl_red) wave_max_to(A.rk.cfl_red, cfl);
}
__global__ __launch_bounds__(256) void k_rk_update(UpdArgs A)
{
    rk_body(A, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- viscous derivatives ------------------------------------------------------------------
// grad order: u_xi,u_eta,u_zeta, v_xi,v_eta,v_zeta, w_xi,w_eta,w_zeta, T_xi,T_eta,T_zeta.
// Direction d is evaluated where the node's d index is interior (all transverse nodes).
__device__ __forceinline__ void grad_body(const CFields5 &pr, const gcdp T, const Fields12 &gr,
                                          const Geom &gm, const Consts &k, const long t)
{
    if (t >= gm.ntot) return;
    int idx[3];
    const long n = padded_node(gm, t, idx);
    const int half = gm.g;
    const int nl = 2 * gm.g + 1;   // S + 2 points
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d >= gm.dims) continue;
        if (idx[d] < gm.lo[d] || idx[d] >= gm.lo[d] + gm.n[d]) continue;
        const long st = gm.st[d];
        double du = 0.0, dv = 0.0, dw = 0.0, dT = 0.0;
        for (int l = 0; l < nl; l++) {
            const long m = n + (long)(l - half) * st;
            const double c = k.zd_der[l];
            du += c * pr.f[1][m];
            dv += c * pr.f[2][m];
            dT += c * T[m];
            if (gm.dims == 3) dw += c * pr.f[3][m];
        }
        gr.f[0 + d][n] = du;
        gr.f[3 + d][n] = dv;
        if (gm.dims == 3) gr.f[6 + d][n] = dw;
        gr.f[9 + d][n] = dT;
    }
}
__global__ void k_visc_grad(CFields5 pr, const double *T, Fields12 gr, Geom gm, Consts k)
{
    grad_body(pr, T, gr, gm, k, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// 3D nodal derivatives, marching along i.  A block owns a GK x GJ column of (j, k) nodes over the
// full padded extent and walks a chunk of i planes: every plane is read from HBM once per chunk
// into LDS with a g-wide halo; the j and k derivatives come from there, the i derivative from a
// register window of 2g+1 plane centres per thread.  Outputs as k_visc_grad; the T
// derivatives only when `tgrad` (the diagonal-metric viscous sweeps never read them).
//
// VX (3D diagonal metric): the viscous xi direction rides along (CalcViscidFluxesInXiDirection-
// Directly, src/ZD.c:47-260).  The march already holds, per column, the u, v, w window along i and
// each plane's eta and zeta derivatives; with T, mu and the metric of the column read as the planes
// pass, face i+1/2 follows at plane i+g from rings of the last 2g planes, and node i adds its xi
// viscous terms to Q.  The face and node arithmetic is the xi sweep's (k_visc_sweep<..,0,..>) term
// for term, so Q gets the same bits; u_eta and u_zeta, which only the xi direction reads, are not
// stored.
constexpr int GK = 64, GJ = 16;
struct GradArgs {
    CFields5 pr;
    gcdp T;
    Fields12 gr;
    Geom gm;
    Consts k;
    int ci, tgrad;
    int tg[4];                    // tile grid (k columns, j columns, i chunks) of a single-case launch
    // VX: Q (after the WENO sweeps), mu, the diagonal metric and J; the isothermal upper xi face
    Fields5 q;
    gcdp mu, m0, m4, m8, J;
    int iso_hi;
    double gam_iso;
};
template <int S, bool TG, bool VX = false, int GJ = ::GJ, int GK = ::GK>
__device__ __forceinline__ void march_body(const GradArgs &a, const int bx, const int by, const int bz)
{
    const CFields5 &pr = a.pr;
    const gcdp T = a.T;
    const Fields12 &gr = a.gr;
    const Geom &gm = a.gm;
    const Consts &k = a.k;
    const int ci = a.ci;
    constexpr int NF = TG ? 4 : 3;             // u, v, w (+ T for its derivatives)
    constexpr int G = (S + 1) / 2, NL = 2 * G + 1;
    constexpr int LK = GK + 2 * G, LJ = GJ + 2 * G, LP = LK * LJ;
    __shared__ double sm[NF * LP];
    const int tx = threadIdx.x % GK, ty = threadIdx.x / GK;
    const int k0 = bx * GK, j0 = by * GJ;
    const int kk = k0 + tx, jj = j0 + ty;
    const bool mine = (kk < gm.ng[2]) && (jj < gm.ng[1]);
    const int i0 = bz * ci, i1 = min(i0 + ci, gm.ng[0]);
    const gcdp src[4] = {pr.f[1], pr.f[2], pr.f[3], T};
    const long col = (long)jj * gm.st[1] + kk;
    constexpr int NT = GK * GJ, NST = (LP + NT - 1) / NT;   // staged plane elements per thread
    const int c0 = (ty + G) * LK + (tx + G);                  // this thread's node in a staged plane
    const bool jin = mine && jj >= gm.lo[1] && jj < gm.lo[1] + gm.n[1];
    const bool kin = mine && kk >= gm.lo[2] && kk < gm.lo[2] + gm.n[2];
    // Planes i0-G .. i1+G-1 pass through LDS once each (with their j/k halo): plane p gives its
    // j and k derivatives when p lies in the chunk, and its centre value enters the register
    // window of 2g+1 planes, from which the i derivative of plane p-G follows.  Planes outside
    // the padded extent are clamped; they only reach windows of ghost planes, whose i derivatives
    // are not written.
    double nxt[NST][NF];
    auto fetch = [&](const int p) {
        const long pbase = (long)max(0, min(p, gm.ng[0] - 1)) * gm.st[0];
#pragma unroll
        for (int s = 0; s < NST; s++) {
            const int e = threadIdx.x + s * NT;
            const int lj = e / LK, lk = e - lj * LK;
            const int j = j0 - G + lj, kq = k0 - G + lk;
            const bool in = (e < LP) && (j >= 0) && (j < gm.ng[1]) && (kq >= 0) && (kq < gm.ng[2]);
            const long m = pbase + (long)j * gm.st[1] + kq;
#pragma unroll
            for (int f = 0; f < NF; f++) nxt[s][f] = in ? src[f][m] : 0.0;
        }
    };
    double win[NF][NL];
#pragma unroll
    for (int l = 0; l < NL; l++)
#pragma unroll
        for (int f = 0; f < NF; f++) win[f][l] = 0.0;
    // VX rings over the last 2g planes (index 2G-1 = plane p): T and the cross derivatives
    // v_eta, u_eta, w_zeta, u_zeta (the xi sweep's cross slots 0..3); mu and metric of planes
    // i and i+1 of the face; the previous face's flux
    constexpr int NR = VX ? 2 * G : 1;
    double tw[NR], cw[4][NR];
    double mv0[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, mv1[5] = {0.0, 0.0, 0.0, 0.0, 0.0};   // mu, m0, m4, m8, J
    double fprev[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (VX) {
#pragma unroll
        for (int l = 0; l < NR; l++) {
            tw[l] = 0.0;
#pragma unroll
            for (int c = 0; c < 4; c++) cw[c][l] = 0.0;
        }
    }
    const bool vcol = VX && jin && kin;
    const int ilo = gm.lo[0], ihi = gm.lo[0] + gm.n[0];
    // the next plane is prefetched where the registers allow it (WENO5 order, no T derivatives)
    constexpr bool PF = (S == 5) && !TG;
    const int pend = i1 + G;
    if constexpr (PF) fetch(i0 - G);
    for (int p = i0 - G; p < pend; p++) {
        if constexpr (PF) {
#pragma unroll
            for (int s = 0; s < NST; s++) {
                const int e = threadIdx.x + s * NT;
                if (e < LP) {
#pragma unroll
                    for (int f = 0; f < NF; f++) sm[f * LP + e] = nxt[s][f];
                }
            }
        } else {
            const long pbase = (long)max(0, min(p, gm.ng[0] - 1)) * gm.st[0];
            for (int e = threadIdx.x; e < LP; e += NT) {
                const int lj = e / LK, lk = e - lj * LK;
                const int j = j0 - G + lj, kq = k0 - G + lk;
                const bool in = (j >= 0) && (j < gm.ng[1]) && (kq >= 0) && (kq < gm.ng[2]);
                const long m = pbase + (long)j * gm.st[1] + kq;
#pragma unroll
                for (int f = 0; f < NF; f++) sm[f * LP + e] = in ? src[f][m] : 0.0;
            }
        }
        __syncthreads();
        if constexpr (PF) {
            if (p + 1 < pend) fetch(p + 1);
        }
#pragma unroll
        for (int l = 0; l < NL - 1; l++)
#pragma unroll
            for (int f = 0; f < NF; f++) win[f][l] = win[f][l + 1];
#pragma unroll
        for (int f = 0; f < NF; f++) win[f][NL - 1] = sm[f * LP + c0];
        if (mine) {
            const bool chunk = (p >= i0 && p < i1);
            const long n = (long)p * gm.st[0] + col;
            if (jin && (chunk || vcol)) {
                double d[NF] = {};
#pragma unroll
                for (int l = 0; l < NL; l++)
#pragma unroll
                    for (int f = 0; f < NF; f++) d[f] += k.zd_der[l] * sm[f * LP + c0 + (l - G) * LK];
                if (chunk) {
                    if (!VX) gr.f[1][n] = d[0];
                    gr.f[4][n] = d[1]; gr.f[7][n] = d[2];
                    if constexpr (TG) gr.f[10][n] = d[3];
                }
                if constexpr (VX) {
#pragma unroll
                    for (int l = 0; l < NR - 1; l++) { cw[0][l] = cw[0][l + 1]; cw[1][l] = cw[1][l + 1]; }
                    cw[0][NR - 1] = d[1];   // v_eta
                    cw[1][NR - 1] = d[0];   // u_eta
                }
            }
            if (kin && (chunk || vcol)) {
                double d[NF] = {};
#pragma unroll
                for (int l = 0; l < NL; l++)
#pragma unroll
                    for (int f = 0; f < NF; f++) d[f] += k.zd_der[l] * sm[f * LP + c0 + (l - G)];
                if (chunk) {
                    if (!VX) gr.f[2][n] = d[0];
                    gr.f[5][n] = d[1]; gr.f[8][n] = d[2];
                    if constexpr (TG) gr.f[11][n] = d[3];
                }
                if constexpr (VX) {
#pragma unroll
                    for (int l = 0; l < NR - 1; l++) { cw[2][l] = cw[2][l + 1]; cw[3][l] = cw[3][l + 1]; }
                    cw[2][NR - 1] = d[2];   // w_zeta
                    cw[3][NR - 1] = d[0];   // u_zeta
                }
            }
            const int i = p - G;
            double gx[NF] = {};         // i derivatives of node i
            const bool iout = (i >= i0 && i >= gm.lo[0] && i < gm.lo[0] + gm.n[0]);
            if (iout) {
                const long ni = (long)i * gm.st[0] + col;
#pragma unroll
                for (int l = 0; l < NL; l++)
#pragma unroll
                    for (int f = 0; f < NF; f++) gx[f] += k.zd_der[l] * win[f][l];
                gr.f[0][ni] = gx[0]; gr.f[3][ni] = gx[1]; gr.f[6][ni] = gx[2];
                if constexpr (TG) gr.f[9][ni] = gx[3];
            }
            if constexpr (VX) {
                if (vcol) {
                    // T of plane p; mu and metric of plane i+1 (the face's upper node).  Loading
                    // these and Q one plane ahead was measured slower at 256^3 (0.95 -> 1.03 ms)
                    const long pc = (long)max(0, min(p, gm.ng[0] - 1)) * gm.st[0] + col;
#pragma unroll
                    for (int l = 0; l < NR - 1; l++) tw[l] = tw[l + 1];
                    tw[NR - 1] = T[pc];
                    const long pm = (long)max(0, min(i + 1, gm.ng[0] - 1)) * gm.st[0] + col;
#pragma unroll
                    for (int c = 0; c < 5; c++) mv0[c] = mv1[c];
                    mv1[0] = a.mu[pm]; mv1[1] = a.m0[pm]; mv1[2] = a.m4[pm]; mv1[3] = a.m8[pm]; mv1[4] = a.J[pm];
                    if (i >= max(i0, ilo) - 1 && i < min(i1, ihi)) {
                        // face i+1/2 from planes i-G+1 .. i+G: window slots 1..2G, ring slots 0..2G-1
                        double own4[4] = {0.0, 0.0, 0.0, 0.0}, cr[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                        for (int l = 0; l <= S; l++) {
                            const double cm = k.zd_mid[l], cint = k.zd_int[l];
                            own4[0] += cm * win[0][l + 1];
                            own4[1] += cm * win[1][l + 1];
                            own4[2] += cm * win[2][l + 1];
                            own4[3] += cm * tw[l];
#pragma unroll
                            for (int c = 0; c < 4; c++) cr[c] += cint * cw[c][l];
                        }
                        double dd[12];
#pragma unroll
                        for (int s2 = 0; s2 < 12; s2++) dd[s2] = 0.0;
                        dd[0 * 3 + 0] = own4[0]; dd[1 * 3 + 0] = own4[1]; dd[2 * 3 + 0] = own4[2]; dd[9 + 0] = own4[3];
                        dd[1 * 3 + 1] = cr[0]; dd[0 * 3 + 1] = cr[1];
                        dd[2 * 3 + 2] = cr[2]; dd[0 * 3 + 2] = cr[3];
                        const double g0[3] = {mv0[1], mv0[2], mv0[3]}, g1[3] = {mv1[1], mv1[2], mv1[3]};
                        double F[4];
                        face_flux_diag<0>(dd, g0, mv0[4], g1, mv1[4], mv0[0], mv1[0], F);
                        if (i >= max(i0, ilo)) {
                            // node i: faces i-1/2 (fprev) and i+1/2
                            const long ni = (long)i * gm.st[0] + col;
                            double dF[4];
#pragma unroll
                            for (int m = 0; m < 4; m++) dF[m] = F[m] - fprev[m];
                            double g[9];
                            g[0] = gx[0]; g[3] = gx[1]; g[6] = gx[2];
                            // cross derivatives of node i: ring slot G-1 (plane i)
                            g[1 * 3 + 1] = cw[0][G - 1]; g[0 * 3 + 1] = cw[1][G - 1];
                            g[2 * 3 + 2] = cw[2][G - 1]; g[0 * 3 + 2] = cw[3][G - 1];
                            g[1 * 3 + 2] = 0.0; g[2 * 3 + 1] = 0.0;
                            const double gmv[3] = {mv0[1], mv0[2], mv0[3]};
                            const double prod = node_prod_diag<0>(g, gmv, mv0[4], mv0[0]);
                            const double work = win[0][G] * dF[0] + win[1][G] * dF[1] + win[2][G] * dF[2];
                            double q0[5];
#pragma unroll
                            for (int m = 1; m < 5; m++) q0[m] = a.q.f[m][ni];
                            a.q.f[1][ni] = q0[1] + k.psi * dF[0];
                            a.q.f[2][ni] = q0[2] + k.psi * dF[1];
                            a.q.f[3][ni] = q0[3] + k.psi * dF[2];
                            const double gh = (a.iso_hi && i == ihi - 1) ? a.gam_iso : k.gam_heat;
                            a.q.f[4][ni] = q0[4] + (gh * dF[3] + k.psi * (work + prod));
                        }
#pragma unroll
                        for (int m = 0; m < 4; m++) fprev[m] = F[m];
                    }
                }
            }
        }
        __syncthreads();
    }
}
template <int S, bool TG>
__global__ __launch_bounds__(GK * GJ) void k_visc_grad_march(GradArgs a)
{
    int bx, by, bz;
    sweep_tile(blockIdx.x, a.tg, bx, by, bz);
    march_body<S, TG>(a, bx, by, bz);
}
// the march with the viscous xi direction (3D diagonal metric): 64 x 8 columns, whose 512 threads
// hold the rings in registers (201 VGPRs, two waves per SIMD).  Measured against it at 256^3: 64 x
// 16 columns spill 160 VGPRs, and 64 x 4 held to three waves per SIMD spills 60 and is slower than
// the fused zeta kernel.
constexpr int GKX = 64, GJX = 8;
template <int S>
__global__ __launch_bounds__(GKX * GJX) void k_visc_grad_march_xi(GradArgs a)
{
    int bx, by, bz;
    sweep_tile(blockIdx.x, a.tg, bx, by, bz);
    march_body<S, false, true, GJX, GKX>(a, bx, by, bz);
}

// ---- batched launches ----------------------------------------------------------------------------
// One launch runs a kernel class over many independent cases.  The launch's work list holds one
// descriptor per block of every case: (case, tile x, tile y, tile z), or (case, block, 0, 0) for
// the node-parallel kernels, precomputed on the host so that a block starts from one scalar load
// and no integer division.  Blocks walk the list (one item per block, or with a grid stride when
// the launch is sized to the device: persistent blocks).
// The per-case records and the work list are read-only for the launch and are addressed in the
// constant address space, as kernel arguments are: the compiler reads them with scalar loads and,
// since nothing can write them, keeps the values across the LDS fences of the sweeps instead of
// reloading them (a generic record pointer would be reloaded after every fence).
#define KCONST __attribute__((address_space(4)))
#define BATCH_PARAMS(T) const KCONST T *__restrict__ A, const KCONST int4 *__restrict__ work, const int total
template <class T>
__device__ __forceinline__ const T &rec(const KCONST T *A, const int i)
{
    return *(const T *)&A[i];
}

// P = false: one block per work item (the kernel then needs exactly the registers of its
// single-case twin); P = true: persistent blocks walking the list with a grid stride.
template <bool P, class F>
__device__ __forceinline__ void batch_walk(const int total, F &&f)
{
    if constexpr (!P) {
        if ((int)blockIdx.x < total) f((int)blockIdx.x);
    } else {
        for (int b = blockIdx.x; b < total; b += gridDim.x) {
            f(b);
            __syncthreads();
        }
    }
}

struct GhostArgs { Fields10 f; int nf; Geom gm; };
struct ValuesArgs { CFields5 pr; gdp T, mu, P, C; Geom gm; Consts k; };
template <bool P>
__global__ __launch_bounds__(256) void k_ghost_batch(BATCH_PARAMS(GhostArgs))
{
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        const GhostArgs &a = rec(A, w.x);
        ghost_body(a.f, a.nf, a.gm, (long)w.y * 256 + threadIdx.x);
    });
}
template <bool P>
__global__ __launch_bounds__(256) void k_values_batch(BATCH_PARAMS(ValuesArgs))
{
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        const ValuesArgs &a = rec(A, w.x);
        values_body(a.pr, a.T, a.mu, a.P, a.C, a.gm, a.k, (long)w.y * 256 + threadIdx.x);
    });
}
struct GhostValuesArgs { Fields5 pr; gdp T, mu, P, C; Geom gm; Consts k; };
template <bool P>
__global__ __launch_bounds__(256) void k_ghost_values_batch(BATCH_PARAMS(GhostValuesArgs))
{
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        const GhostValuesArgs &a = rec(A, w.x);
        ghost_values_body(a.pr, a.T, a.mu, a.P, a.C, a.gm, a.k, (long)w.y * 256 + threadIdx.x);
    });
}
template <bool P>
__global__ __launch_bounds__(256) void k_rk_batch(BATCH_PARAMS(UpdArgs))
{
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        rk_body(rec(A, w.x), (long)w.y * 256 + threadIdx.x);
    });
}
template <bool P>
__global__ __launch_bounds__(256) void k_face_batch(BATCH_PARAMS(BcArgs))
{
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        face_body(rec(A, w.x), (long)w.y * 256 + threadIdx.x);
    });
}
template <bool P>
__global__ __launch_bounds__(256) void k_face_save_batch(BATCH_PARAMS(BcArgs))
{
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        face_save_body(rec(A, w.x), (long)w.y * 256 + threadIdx.x);
    });
}
struct CarryArgs { CFields5 from; Fields5 to; Geom gm; };
template <bool P>
__global__ __launch_bounds__(256) void k_carry_batch(BATCH_PARAMS(CarryArgs))
{
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        const CarryArgs &a = rec(A, w.x);
        carry_body(a.from, a.to, a.gm, (long)w.y * 256 + threadIdx.x);
    });
}
template <bool P>
__global__ __launch_bounds__(256) void k_grad2d_batch(BATCH_PARAMS(GradArgs))
{
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        const GradArgs &a = rec(A, w.x);
        grad_body(a.pr, a.T, a.gr, a.gm, a.k, (long)w.y * 256 + threadIdx.x);
    });
}

template <int S, bool P, bool TG>
__global__ __launch_bounds__(GK * GJ) void k_march_batch(BATCH_PARAMS(GradArgs))
{
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        march_body<S, TG>(rec(A, w.x), w.y, w.z, w.w);
    });
}

// derivative march kernel for an order, with or without T derivatives (the general metric), as a
// single-case launch or a batched one (persistent or not)
static const void *march_kernel(int S, bool tgrad, bool batch, bool persistent = false)
{
    if (!batch) {
        if (S == 5) return tgrad ? (const void *)&k_visc_grad_march<5, true> : (const void *)&k_visc_grad_march<5, false>;
        return tgrad ? (const void *)&k_visc_grad_march<9, true> : (const void *)&k_visc_grad_march<9, false>;
    }
    if (S == 5) {
        if (tgrad) return persistent ? (const void *)&k_march_batch<5, true, true> : (const void *)&k_march_batch<5, false, true>;
        return persistent ? (const void *)&k_march_batch<5, true, false> : (const void *)&k_march_batch<5, false, false>;
    }
    if (tgrad) return persistent ? (const void *)&k_march_batch<9, true, true> : (const void *)&k_march_batch<9, false, true>;
    return persistent ? (const void *)&k_march_batch<9, true, false> : (const void *)&k_march_batch<9, false, false>;
}

template <int S, int TA, int TDI, int AX, bool MIX, bool P>
__global__ __launch_bounds__(TA * TDI, (weno_min_waves<S, AX, MIX>())) void k_weno_batch(BATCH_PARAMS(WenoArgs))
{
    __shared__ double lds[weno_lds<S, TA, TDI, AX, MIX>()];
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        weno_body<S, TA, TDI, AX, MIX>(rec(A, w.x), w.y, w.z, w.w, lds);
    });
}

template <int S, int DIR, int TA, int TDI, bool DIAG, bool P>
__global__ __launch_bounds__(TA * TDI) void k_visc_batch(BATCH_PARAMS(ViscArgs))
{
    __shared__ double lds[visc_lds<S, DIR, TA, TDI, DIAG>()];
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        visc_body<S, DIR, TA, TDI, DIAG>(rec(A, w.x), w.y, w.z, w.w, lds);
    });
}
template <int S, bool MIX, bool P>
__global__ __launch_bounds__(256, ((S == 5 && !MIX) ? 4 : 1)) void k_zeta_batch(BATCH_PARAMS(ZetaArgs))
{
    __shared__ double lds[zeta_lds<S, MIX>()];
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        zeta_body<S, MIX>(rec(A, w.x), w.y, w.z, w.w, lds);
    });
}

template <bool MIX, bool P>
__global__ __launch_bounds__(256, (MIX ? 1 : ETA_WAVES)) void k_eta_batch(BATCH_PARAMS(EtaArgs))
{
    __shared__ double lds[eta_lds<MIX>()];
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        eta_body<MIX>(rec(A, w.x), w.y, w.z, w.w, lds);
    });
}
template <bool P>
__global__ __launch_bounds__(GKX * GJX) void k_march_xi_batch(BATCH_PARAMS(GradArgs))
{
    batch_walk<P>(total, [&](const int b) {
        const int4 w = rec(work, b);
        march_body<5, false, true, GJX, GKX>(rec(A, w.x), w.y, w.z, w.w);
    });
}

// Q += J * src (manufactured-solution source) for the stage residual returned by shx_rhs.
__global__ void k_add_source(Fields5 q, CFields5 src, const double *J, Geom gm)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    if (t >= nint) return;
    const int kk = (int)(t % gm.n[2]);
    const long r = t / gm.n[2];
    const int jj = (int)(r % gm.n[1]);
    const int ii = (int)(r / gm.n[1]);
    const long n = (long)(ii + gm.lo[0]) * gm.st[0] + (long)(jj + gm.lo[1]) * gm.st[1] + kk + gm.lo[2];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        if (m == 3 && gm.dims == 2) continue;
        q.f[m][n] += J[n] * src.f[m][n];
    }
}

// ---- metric --------------------------------------------------------------------------------
// CreateMetric (src/Functions.c:6388-6509) on the extended coordinates.
__global__ void k_metric(const double *xe, const double *ye, const double *ze, Fields10 met, Geom gm,
                         int diag3d)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= gm.ntot) return;
    int idx[3];
    const long n = padded_node(gm, t, idx);
    const int i = idx[0], j = idx[1], k = idx[2];
    const int ne1 = gm.ng[1] + 2, ne2 = (gm.dims == 3) ? gm.ng[2] + 2 : 1;
    const int ke = (gm.dims == 3) ? k + 1 : 0;
    const long ce = ((long)(i + 1) * ne1 + (j + 1)) * ne2 + ke;
    const long sI = (long)ne1 * ne2, sJ = ne2;
    double x_xi, y_xi, z_xi, x_eta, y_eta, z_eta, x_zeta, y_zeta, z_zeta;
    x_xi = (-0.5 * xe[ce - sI] + 0.5 * xe[ce + sI]);
    y_xi = (-0.5 * ye[ce - sI] + 0.5 * ye[ce + sI]);
    x_eta = (-0.5 * xe[ce - sJ] + 0.5 * xe[ce + sJ]);
    y_eta = (-0.5 * ye[ce - sJ] + 0.5 * ye[ce + sJ]);
    if (gm.dims == 2) {
        z_xi = 0.0; z_eta = 0.0; x_zeta = 0.0; y_zeta = 0.0; z_zeta = 1.0;
    } else {
        z_xi = (-0.5 * ze[ce - sI] + 0.5 * ze[ce + sI]);
        z_eta = (-0.5 * ze[ce - sJ] + 0.5 * ze[ce + sJ]);
        x_zeta = (-0.5 * xe[ce - 1] + 0.5 * xe[ce + 1]);
        y_zeta = (-0.5 * ye[ce - 1] + 0.5 * ye[ce + 1]);
        z_zeta = (-0.5 * ze[ce - 1] + 0.5 * ze[ce + 1]);
        if (diag3d) { y_xi = 0.0; z_xi = 0.0; x_eta = 0.0; z_eta = 0.0; x_zeta = 0.0; y_zeta = 0.0; }
    }
    const double J = x_xi * y_eta * z_zeta + x_eta * y_zeta * z_xi + x_zeta * y_xi * z_eta
                   - x_zeta * y_eta * z_xi - x_eta * y_xi * z_zeta - x_xi * y_zeta * z_eta;
    met.f[9][n] = J;
    met.f[0][n] = (y_eta * z_zeta - y_zeta * z_eta) / J;
    met.f[1][n] = (x_zeta * z_eta - x_eta * z_zeta) / J;
    met.f[2][n] = (x_eta * y_zeta - x_zeta * y_eta) / J;
    met.f[3][n] = (y_zeta * z_xi - y_xi * z_zeta) / J;
    met.f[4][n] = (x_xi * z_zeta - x_zeta * z_xi) / J;
    met.f[5][n] = (x_zeta * y_xi - x_xi * y_zeta) / J;
    met.f[6][n] = (y_xi * z_eta - y_eta * z_xi) / J;
    met.f[7][n] = (x_eta * z_xi - x_xi * z_eta) / J;
    met.f[8][n] = (x_xi * y_eta - x_eta * y_xi) / J;
}

// ---- reductions ----------------------------------------------------------------------------

// CFL spectral-radius maximum over interior nodes; positive doubles order like their bit
// patterns, so the block maxima combine with an unsigned 64-bit atomic max.
__global__ void k_cfl_max(CFields5 pr, CFields10 met, Geom gm, Consts k, unsigned long long *out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    double v = 0.0;
    if (t < nint) {
        const int kk = (int)(t % gm.n[2]);
        const long r = t / gm.n[2];
        const int jj = (int)(r % gm.n[1]);
        const int ii = (int)(r / gm.n[1]);
        const long n = (long)(ii + gm.lo[0]) * gm.st[0] + (long)(jj + gm.lo[1]) * gm.st[1] + kk + gm.lo[2];
        double m9[9];
#pragma unroll
        for (int c = 0; c < 9; c++) m9[c] = met.f[c][n];
        v = node_cfl_sum(k.gamma, k.upsilon, gm.dims, pr.f[0][n], pr.f[1][n], pr.f[2][n], pr.f[3][n], pr.f[4][n], m9);
    }
    v = wave_max(v);
    __shared__ double red[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int s = 0; s < (int)(blockDim.x >> 6); s++) m = fmax(m, red[s]);
        atomicMax(out + (blockIdx.x & (RED_SLOTS - 1)), (unsigned long long)__double_as_longlong(m));
    }
}

// checkNAN (src/Functions.c:5419-5467): count and overwrite non-finite interior rho (and the
// stored p, where it is kept).
__global__ void k_check_nan(Fields5 pr, gdp P, Geom gm, int *count)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    int bad = 0;
    if (t < nint) {
        const int kk = (int)(t % gm.n[2]);
        const long r = t / gm.n[2];
        const int jj = (int)(r % gm.n[1]);
        const int ii = (int)(r / gm.n[1]);
        const long n = (long)(ii + gm.lo[0]) * gm.st[0] + (long)(jj + gm.lo[1]) * gm.st[1] + kk + gm.lo[2];
        const double rho = pr.f[0][n];
        if (!isfinite(rho)) {
            bad = 1;
            pr.f[0][n] = 1000000; pr.f[1][n] = 1000000; pr.f[2][n] = 1000000; pr.f[3][n] = 1000000;
            if (P) P[n] = 1000000;
        }
    }
    bad = wave_sum(bad);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(count, bad);
}

// ---- post-processing ---------------------------------------------------------------------
// lambda2 of interior node n (CalcLambda2, src/Functions.c:5846-6070): velocity derivatives as
// differences of midpoint-interpolated values (the viscous interpolation table), the Cartesian
// gradient through the metric, the middle eigenvalue of S^2 + Omega^2 by the trigonometric
// solution of the characteristic cubic, capped at 0.  The reference's strain tensor carries half
// the derivative on its diagonal (s11 = 0.5 u_x, src/Functions.c:5952-5957); kept.
__device__ double lambda2_node(const CFields5 &pr, const CFields10 &met, const Geom &gm, const Consts &k,
                               const int S, const long n)
{
    const int half = (S + 1) / 2;
    double dv[3][3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double lo[3] = {0., 0., 0.}, hi[3] = {0., 0., 0.};
        for (int m = 0; m <= S; m++) {
            const long i0 = n + (long)(m - half) * gm.st[d], i1 = i0 + gm.st[d];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                lo[q] += k.zd_int[m] * pr.f[1 + q][i0];
                hi[q] += k.zd_int[m] * pr.f[1 + q][i1];
            }
        }
#pragma unroll
        for (int q = 0; q < 3; q++) dv[q][d] = hi[q] - lo[q];
    }
    const double xx = met.f[0][n], xy = met.f[1][n], xz = met.f[2][n], ex = met.f[3][n], ey = met.f[4][n];
    const double ez = met.f[5][n], zx = met.f[6][n], zy = met.f[7][n], zz = met.f[8][n];
    const double udx = dv[0][0] * xx + dv[0][1] * ex + dv[0][2] * zx, udy = dv[0][0] * xy + dv[0][1] * ey + dv[0][2] * zy;
    const double udz = dv[0][0] * xz + dv[0][1] * ez + dv[0][2] * zz;
    const double vdx = dv[1][0] * xx + dv[1][1] * ex + dv[1][2] * zx, vdy = dv[1][0] * xy + dv[1][1] * ey + dv[1][2] * zy;
    const double vdz = dv[1][0] * xz + dv[1][1] * ez + dv[1][2] * zz;
    const double wdx = dv[2][0] * xx + dv[2][1] * ex + dv[2][2] * zx, wdy = dv[2][0] * xy + dv[2][1] * ey + dv[2][2] * zy;
    const double wdz = dv[2][0] * xz + dv[2][1] * ez + dv[2][2] * zz;
    const double om12 = 0.5 * (udy - vdx), om13 = 0.5 * (udz - wdx), om23 = 0.5 * (vdz - wdy);
    const double s11 = 0.5 * (udx), s12 = 0.5 * (udy + vdx), s13 = 0.5 * (udz + wdx);
    const double s22 = 0.5 * (vdy), s23 = 0.5 * (vdz + wdy), s33 = 0.5 * (wdz);
    const double a11 = (s11 * s11 + s12 * s12 + s13 * s13) + (-om12 * om12 - om13 * om13);
    const double a12 = (s11 * s12 + s12 * s22 + s13 * s23) + (-om13 * om23);
    const double a13 = (s11 * s13 + s12 * s23 + s13 * s33) + (om12 * om23);
    const double a22 = (s12 * s12 + s22 * s22 + s23 * s23) + (-om12 * om12 - om23 * om23);
    const double a23 = (s12 * s13 + s22 * s23 + s23 * s33) + (-om12 * om13);
    const double a33 = (s13 * s13 + s23 * s23 + s33 * s33) + (-om13 * om13 - om23 * om23);
    const double f1 = -a11 - a22 - a33;
    const double f2 = -a23 * a23 - a12 * a12 - a13 * a13 + a11 * a33 + a11 * a22 + a22 * a33;
    const double f3 = -a11 * a22 * a33 + a11 * a23 * a23 + a22 * a13 * a13 + a33 * a12 * a12 - 2. * a12 * a13 * a23;
    const double p = -f1 * f1 / 3. + f2, q = 2. * f1 * f1 * f1 / 27. - f1 * f2 / 3. + f3;
    const double rrr = sqrt(fabs(p) / 3.);
    const double pi = 3.14159265358979323846;
    const double al = acos(q / (2. * rrr * rrr * rrr));
    const double l1 = -2. * rrr * cos(al / 3.) - f1 / 3, l2 = 2. * rrr * cos(pi / 3. - al / 3.) - f1 / 3;
    const double l3 = 2. * rrr * cos(pi / 3. + al / 3.) - f1 / 3;
    double lmax = (l2 >= l1) ? l2 : l1;
    if (l3 >= lmax) lmax = l3;
    double lmin = (l2 <= l1) ? l2 : l1;
    if (l3 <= lmin) lmin = l3;
    double lam = 0.;
    if (l1 == l2 || l2 == l3) lam = l2;
    if (l1 > lmin && l1 < lmax) lam = l1;
    else if (l2 > lmin && l2 < lmax) lam = l2;
    else if (l3 > lmin && l3 < lmax) lam = l3;
    return (lam >= 0.) ? 0. : lam;
}

// CalcValuesForPost (src/Functions.c:3687-3761) on interior nodes: Mach number from the stored
// pressure (U_lastStep's, which the RK update leaves one stage behind), the density "gradient" as
// the root of the sum of squares of every (central difference x metric entry) product, and
// lambda2 (3D).  out: Mach, gradRho, Lambda2 on the padded grid.
__global__ void k_post(CFields5 pr, gcdp P, CFields10 met, Fields5 out, Geom gm, Consts k, int S)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    if (t >= nint) return;
    const int kk = (int)(t % gm.n[2]);
    const long r = t / gm.n[2];
    const int jj = (int)(r % gm.n[1]);
    const int ii = (int)(r / gm.n[1]);
    const long n = (long)(ii + gm.lo[0]) * gm.st[0] + (long)(jj + gm.lo[1]) * gm.st[1] + kk + gm.lo[2];
    const double rho = pr.f[0][n], u = pr.f[1][n], v = pr.f[2][n], w = pr.f[3][n];
    out.f[0][n] = (sqrt(u * u + v * v + w * w)) / sqrt(P[n] / rho) * k.mach;
    const double dxi = (0.5 * pr.f[0][n + gm.st[0]] - 0.5 * pr.f[0][n - gm.st[0]]) / 1.0;
    const double deta = (0.5 * pr.f[0][n + gm.st[1]] - 0.5 * pr.f[0][n - gm.st[1]]) / 1.0;
    auto sq = [](const double x) { return pow(x, 2.0); };
    if (gm.dims == 2) {
        out.f[1][n] = sqrt(sq(dxi * met.f[0][n]) + sq(deta * met.f[3][n]) + sq(dxi * met.f[1][n]) + sq(deta * met.f[4][n]));
    } else {
        const double dze = (0.5 * pr.f[0][n + gm.st[2]] - 0.5 * pr.f[0][n - gm.st[2]]) / 1.0;
        out.f[1][n] = sqrt(sq(dxi * met.f[0][n]) + sq(deta * met.f[3][n]) + sq(dze * met.f[6][n])
                           + sq(dxi * met.f[1][n]) + sq(deta * met.f[4][n]) + sq(dze * met.f[7][n])
                           + sq(dxi * met.f[2][n]) + sq(deta * met.f[5][n]) + sq(dze * met.f[8][n]));
        out.f[2][n] = lambda2_node(pr, met, gm, k, S, n);
    }
}

// WriteValuesFromUToFilm (src/Functions.c:2510-2553): rho, u, v, w, p, gradRho, Mach, Lambda2 of
// the interior nodes into sample `s` of the device-resident film (interior-compact, field-major)
__global__ void k_film(CFields5 pr, gcdp P, CFields5 post, gdp film, long nint, int s, Geom gm)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nint) return;
    const int kk = (int)(t % gm.n[2]);
    const long r = t / gm.n[2];
    const int jj = (int)(r % gm.n[1]);
    const int ii = (int)(r / gm.n[1]);
    const long n = (long)(ii + gm.lo[0]) * gm.st[0] + (long)(jj + gm.lo[1]) * gm.st[1] + kk + gm.lo[2];
    const long base = (long)s * 8 * nint + t;
    film[base + 0 * nint] = pr.f[0][n];
    film[base + 1 * nint] = pr.f[1][n];
    film[base + 2 * nint] = pr.f[2][n];
    film[base + 3 * nint] = pr.f[3][n];
    film[base + 4 * nint] = P[n];
    film[base + 5 * nint] = post.f[1][n];
    film[base + 6 * nint] = post.f[0][n];
    film[base + 7 * nint] = post.f[2][n];
}

// probe histories (src/Functions.c:5381-5417): stored p and u, v, w of each probe node into row
// `row` of the history (field-major: p, u, v, w, each [nprobes][iters])
__global__ void k_probe(CFields5 pr, gcdp P, const long *idx, int np, gdp hist, int row, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const long n = idx[t];
    const long base = (long)t * iters + row, fs = (long)np * iters;
    hist[base] = P[n];
    hist[base + fs] = pr.f[1][n];
    hist[base + 2 * fs] = pr.f[2][n];
    hist[base + 3 * fs] = pr.f[3][n];
}

// ErrorManufacturedSolution (src/ManufacturedSolution.c:745-790): squared deviations of rho and of
// the stored p from the exact values, each divided by the interior node count N, and their largest
// magnitudes, over the interior nodes of the last k plane (the reference's k loop starts at
// kEndReal).  Pass 1: per-block partial sums and maxima; pass 2 (one block): the totals.
__device__ __forceinline__ void block_sum_max(double &s0, double &m0, double &s1, double &m1, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = s0; sh[256 + t] = m0; sh[512 + t] = s1; sh[768 + t] = m1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            sh[t] += sh[t + w];
            sh[256 + t] = fmax(sh[256 + t], sh[256 + t + w]);
            sh[512 + t] += sh[512 + t + w];
            sh[768 + t] = fmax(sh[768 + t], sh[768 + t + w]);
        }
        __syncthreads();
    }
    s0 = sh[0]; m0 = sh[256]; s1 = sh[512]; m1 = sh[768];
}
__global__ __launch_bounds__(256) void k_mms_error(CFields5 pr, gcdp P, gcdp ex_rho, gcdp ex_p, Geom gm,
                                                  double N, gdp part)
{
    __shared__ double sh[1024];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nplane = (long)gm.n[0] * gm.n[1];
    double s0 = 0.0, m0 = 0.0, s1 = 0.0, m1 = 0.0;
    if (t < nplane) {
        const int jj = (int)(t % gm.n[1]), ii = (int)(t / gm.n[1]);
        const long n = (long)(ii + gm.lo[0]) * gm.st[0] + (long)(jj + gm.lo[1]) * gm.st[1] + (gm.lo[2] + gm.n[2] - 1);
        const double dr = ex_rho[n] - pr.f[0][n], dp = ex_p[n] - P[n];
        s0 = (dr * dr) / N; m0 = fabs(dr);
        s1 = (dp * dp) / N; m1 = fabs(dp);
    }
    block_sum_max(s0, m0, s1, m1, sh);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 0] = s0; part[4 * blockIdx.x + 1] = m0;
        part[4 * blockIdx.x + 2] = s1; part[4 * blockIdx.x + 3] = m1;
    }
}
__global__ __launch_bounds__(256) void k_mms_error_total(gcdp part, int nb, gdp out)
{
    __shared__ double sh[1024];
    double s0 = 0.0, m0 = 0.0, s1 = 0.0, m1 = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        s0 += part[4 * b + 0]; m0 = fmax(m0, part[4 * b + 1]);
        s1 += part[4 * b + 2]; m1 = fmax(m1, part[4 * b + 3]);
    }
    block_sum_max(s0, m0, s1, m1, sh);
    if (threadIdx.x == 0) { out[0] = sqrt(s0); out[1] = m0; out[2] = sqrt(s1); out[3] = m1; }
}

// check_Metric (src/Functions.c:6198-6243): count the padded nodes with a negative Jacobian.
__global__ void k_check_metric(gcdp J, Geom gm, int *count)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    if (t < gm.ntot) {
        int idx[3];
        const long n = padded_node(gm, t, idx);
        bad = (J[n] < 0.0) ? 1 : 0;
    }
    bad = wave_sum(bad);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(count, bad);
}

// Plain streaming copy of n doubles (calibration of hardware byte counters).
__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// ----------------------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------------------

static char g_err[512];

static int set_err(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_err(SHX_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                               \
    } while (0)

enum { KT_GHOST, KT_VALUES, KT_WENO0, KT_WENO1, KT_WENO2, KT_GRAD, KT_VISC0, KT_VISC1, KT_VISC2, KT_RK, KT_REDUCE, KT_METRIC, KT_FACES, KT_ZETA, KT_GRADXI, KT_ETA };
static const char *g_knames[SHX_NUM_KERNELS] = {
    "ghost_periodic", "ghost_values", "weno_xi", "weno_eta", "weno_zeta",
    "visc_grad", "visc_xi", "visc_eta", "visc_zeta", "rk_update", "reduce", "metric", "ghost_faces",
    "weno_visc_zeta", "visc_grad_xi", "weno_visc_eta"};

// face kinds whose writers store a p and c that the equation of state does not give from the
// ghost's rho,u,v,w,e (subsonic outflow; the slip wall, whose ghost velocity is not the mirror's)
static bool stores_pc(const int kind)
{
    return kind == SHX_BC_OUTFLOW_NORMAL || kind == SHX_BC_OUTFLOW_RUDY || kind == SHX_BC_WALL_SLIP;
}
static bool is_wall(const int kind)
{
    return kind == SHX_BC_WALL_SLIP || kind == SHX_BC_WALL_NOSLIP || kind == SHX_BC_WALL_ISOTHERMAL;
}
// faces whose ghost metric is the mirror's (mirrorGhostCellsMetric, src/Functions.c:6849-7014) and
// whose writers cover the interior tangential range only
static bool mirrors_metric(const int kind) { return is_wall(kind) || kind == SHX_BC_FARFIELD; }
// BC_Corrector of a face's ghost nodes: slip walls reverse the momentum component along the
// face's index direction (src/BC.c:988-1172), no-slip walls all three (src/BC.c:2657-2686)
static void wall_corrector(const int kind, const int dir, double cr[3])
{
    for (int q = 0; q < 3; q++) cr[q] = 1.0;
    if (kind == SHX_BC_WALL_SLIP) cr[dir] = -1.0;
    if (kind == SHX_BC_WALL_NOSLIP || kind == SHX_BC_WALL_ISOTHERMAL) cr[0] = cr[1] = cr[2] = -1.0;
}

struct shx_case {
    shx_desc d;
    Geom gm;
    Consts k;
    double rk_a[4], rk_b[4], rk_q[4], rk_f[4], rk_m[4];
    hipStream_t stream;
    double *prim[3][5];    // three state buffers; cur = U_lastStep
    int cur;
    double *q[5], *qsum[5], *met[10], *grad[12], *src[5], *T, *mu;
    double *bs[6];         // exact state for manufactured faces
    double *P, *C;         // stored p and c (pc: some face is a subsonic outflow)
    double *bk[2][5];      // shx_simulate's two backups of the state (and of P), allocated on use
    double *bkP[2];
    double *post[3];       // Mach, gradRho, Lambda2 (padded), allocated on use
    double *film;          // device-resident samples: 8 interior fields each
    int last_in;           // buffer of the last stage's input of the last step (-1: none)
    long *d_probes;        // probe histories of shx_simulate
    double *hist;
    int hist_np, hist_iters;
    std::vector<double> hist_time;
    int film_cap, film_n;
    std::vector<double> film_time;
    double *PL, *RL, *UL, *C0;   // Rudy faces: step-start ghost p, rho, u; uploaded sound speed
    int pc, rudy;
    int has_src, has_metric, has_state, has_bs;
    int faces;             // number of non-periodic faces
    int needs_bs;          // some face is SHX_BC_MANUFACTURED
    int diag;              // 3D metric is diagonal (only xi_x, eta_y, zeta_z and J non-zero)
    int wtile[3];          // WENO tile variant per direction (set by shx_tune)
    unsigned long long *d_red;
    int *d_cnt;
    int *d_nan;            // running count of the fused checkNAN (shx_run, shx_run_cfl, shx_simulate)
    int nan_seen;          // the count last read
    struct HostIo { unsigned long long red[RED_SLOTS]; int nan; int pad; double mms[4]; } *h_io;   // pinned landing slots
    int dev;               // the device current at shx_create (the stream's)
    int st_in, st_out, st_bx, st_by;   // buffers of the step in progress (step_begin .. step_end)
    int iface;             // faces of kind SHX_BC_INTERFACE
    shx_blocks *blocks = nullptr;      // the multi-block case that holds this block
    int timing;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kind;
    size_t ev_used;
    double kt_ms[SHX_NUM_KERNELS];
    int kt_n[SHX_NUM_KERNELS];
    // the batch that holds this case (its captured step graph uses this case's state buffers as
    // they were at shx_batch_create); stepping the case alone while it is set would leave the
    // batch advancing a stale buffer
    shx_batch *batch = nullptr;
    // pinned staging of host transfers (allocated on the first upload or download): two buffers
    // of stage_bytes, each with the event of the copy that last used it
    double *stage_h[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    size_t stage_bytes = 0;
};

extern "C" const char *shx_last_error(void) { return g_err; }

extern "C" int shx_device_count(int *count)
{
    if (!count) return set_err(SHX_ERR_ARG, "null count");
    HIPCHK(hipGetDeviceCount(count));
    return SHX_OK;
}

// event bracketing for per-kernel timing
static int tbeg(shx_case *c, int kind)
{
    if (!c->timing) return SHX_OK;
    if (c->ev_used + 2 > c->ev_pool.size()) {
        for (int s = 0; s < 256; s++) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            c->ev_pool.push_back(e);
            c->ev_kind.push_back(0);
        }
    }
    c->ev_kind[c->ev_used] = kind;
    HIPCHK(hipEventRecord(c->ev_pool[c->ev_used], c->stream));
    c->ev_used++;
    return SHX_OK;
}
static int tend(shx_case *c)
{
    if (!c->timing) return SHX_OK;
    c->ev_kind[c->ev_used] = -1;
    HIPCHK(hipEventRecord(c->ev_pool[c->ev_used], c->stream));
    c->ev_used++;
    return SHX_OK;
}
static int drain_events(shx_case *c)
{
    if (c->ev_used == 0) return SHX_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t s = 0; s + 1 < c->ev_used; s += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev_pool[s], c->ev_pool[s + 1]));
        c->kt_ms[c->ev_kind[s]] += ms;
        c->kt_n[c->ev_kind[s]] += 1;
    }
    c->ev_used = 0;
    return SHX_OK;
}
#define LAUNCH_CHECK() HIPCHK(hipGetLastError())

static inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

extern "C" int shx_create(const shx_desc *desc, shx_case **out)
{
    if (!desc || !out) return set_err(SHX_ERR_ARG, "null argument");
    *out = nullptr;
    const shx_desc &d = *desc;
    if (d.dims != 2 && d.dims != 3) return set_err(SHX_ERR_ARG, "dims must be 2 or 3 (got %d)", d.dims);
    if (d.space_order != 5 && d.space_order != 9)
        return set_err(SHX_ERR_ARG, "space_order must be 5 or 9 (got %d)", d.space_order);
    if (d.time_order != 3 && d.time_order != 4)
        return set_err(SHX_ERR_ARG, "time_order must be 3 or 4 (got %d)", d.time_order);
    if (d.precision != SHX_FP64 && d.precision != SHX_MIXED)
        return set_err(SHX_ERR_ARG, "precision must be SHX_FP64 or SHX_MIXED (got %d)", d.precision);
    for (int f = 0; f < 2 * d.dims; f++) {
        if (d.bc[f] < SHX_BC_PERIODIC || d.bc[f] > SHX_BC_INTERFACE)
            return set_err(SHX_ERR_ARG, "face %d: unknown boundary kind %d", f, d.bc[f]);
        if ((d.bc[f] == SHX_BC_PERIODIC) != (d.bc[f ^ 1] == SHX_BC_PERIODIC))
            return set_err(SHX_ERR_ARG, "faces %d and %d: a periodic face needs a periodic partner", f & ~1, f | 1);
        if ((d.bc[f] == SHX_BC_INFLOW_SUPERSONIC || d.bc[f] == SHX_BC_OUTFLOW_RIEMANN) &&
            !(d.rho_inflow > 0.0 && d.p_inflow > 0.0))
            return set_err(SHX_ERR_ARG, "face %d: inflow/outflow faces need rho_inflow > 0 and p_inflow > 0", f);
        if ((d.bc[f] == SHX_BC_OUTFLOW_NORMAL || d.bc[f] == SHX_BC_OUTFLOW_RUDY) && !(d.p_out > 0.0 && d.alpha_nonref >= 0.0))
            return set_err(SHX_ERR_ARG, "face %d: subsonic outflow faces need p_out > 0 and alpha_nonref >= 0", f);
        if (d.bc[f] == SHX_BC_WALL_ISOTHERMAL && !(d.T_wall > 0.0))
            return set_err(SHX_ERR_ARG, "face %d: isothermal walls need T_wall > 0", f);
    }
    const int S = d.space_order, g = (S + 1) / 2;
    for (int a = 0; a < d.dims; a++)
        if (d.n[a] < S + 1) return set_err(SHX_ERR_ARG, "n[%d]=%d below the stencil width %d", a, d.n[a], S + 1);
    if (!(d.gamma > 1.0) || !(d.mach > 0.0) || !(d.reynolds > 0.0) || !(d.prandtl > 0.0) || !(d.T0 > 0.0))
        return set_err(SHX_ERR_ARG, "non-physical fluid properties");

    shx_case *c = new shx_case();
    c->last_in = -1;
    c->d = d;
    Geom &gm = c->gm;
    gm.dims = d.dims;
    gm.g = g;
    for (int a = 0; a < 3; a++) {
        const bool live = (a < d.dims);
        gm.n[a] = live ? d.n[a] : 1;
        gm.ng[a] = live ? d.n[a] + S + 1 : 1;
        gm.lo[a] = live ? g : 0;
        gm.per[a] = live ? (d.bc[2 * a] == SHX_BC_PERIODIC) : 1;
        if (live && !gm.per[a]) {
            c->faces += 2;
            if (d.bc[2 * a] == SHX_BC_MANUFACTURED || d.bc[2 * a + 1] == SHX_BC_MANUFACTURED) c->needs_bs = 1;
            for (int u = 0; u < 2; u++) {
                const int kind = d.bc[2 * a + u];
                if (stores_pc(kind)) c->pc = 1;
                if (kind == SHX_BC_OUTFLOW_RUDY) c->rudy = 1;
                if (kind == SHX_BC_INTERFACE) c->iface++;
            }
        }
    }
    // rows of the contiguous axis are padded to 16 doubles (128 B) so tiles start on cache lines
    if (d.dims == 3) {
        const long pitch = ((long)gm.ng[2] + 15) / 16 * 16;
        gm.st[2] = 1; gm.st[1] = pitch; gm.st[0] = (long)gm.ng[1] * pitch;
        gm.nalloc = (long)gm.ng[0] * gm.ng[1] * pitch;
    } else {
        const long pitch = ((long)gm.ng[1] + 15) / 16 * 16;
        gm.st[2] = 1; gm.st[1] = 1; gm.st[0] = pitch;
        gm.nalloc = (long)gm.ng[0] * pitch;
    }
    gm.ntot = (long)gm.ng[0] * gm.ng[1] * gm.ng[2];
    if (gm.nalloc > (1L << 31) - 1) { delete c; return set_err(SHX_ERR_ARG, "block too large for 32-bit tiles"); }

    Consts &k = c->k;
    memset(&k, 0, sizeof k);
    k.gamma = d.gamma;
    k.upsilon = 1.0 / (d.gamma * (d.mach * d.mach));
    k.mach = d.mach;
    k.psi = 1.0 / d.reynolds;
    k.gam_heat = 1.0 / ((d.gamma - 1.0) * pow(d.mach, 2.0) * d.reynolds * d.prandtl);
    k.suth = 100.4 / d.T0;
    k.eps = d.eps_override > 0.0 ? d.eps_override : 10. * pow(DBL_MIN, 0.5);
    k.gm1_ups = (d.gamma - 1.0) / k.upsilon;
    k.mms = d.manufactured ? 1 : 0;
    if (S == 9) {
        const double in9[10] = {1. / 1260., -23. / 2520., 127. / 2520., -473. / 2520., 1627. / 2520.,
                                1627. / 2520., -473. / 2520., 127. / 2520., -23. / 2520., 1. / 1260.};
        const double mi9[10] = {-3. / 9450., 351. / 75600., -2649. / 75600., 15351. / 75600., -110649. / 75600.,
                                110649. / 75600., -15351. / 75600., 2649. / 75600., -351. / 75600., 3. / 9450.};
        const double de9[11] = {-1. / 1260., 25. / 2520., -150. / 2520., 600. / 2520., -2100. / 2520., 0.,
                                2100. / 2520., -600. / 2520., 150. / 2520., -25. / 2520., 1. / 1260.};
        memcpy(k.zd_int, in9, sizeof in9); memcpy(k.zd_mid, mi9, sizeof mi9); memcpy(k.zd_der, de9, sizeof de9);
    } else {
        const double in5[6] = {1. / 60., -8. / 60., 37. / 60., 37. / 60., -8. / 60., 1. / 60.};
        const double mi5[6] = {-1. / 90., 25. / 180., -245. / 180., 245. / 180., -25. / 180., 1. / 90.};
        const double de5[7] = {-1. / 60., 3. / 20., -3. / 4., 0., 3. / 4., -3. / 20., 1. / 60.};
        memcpy(k.zd_int, in5, sizeof in5); memcpy(k.zd_mid, mi5, sizeof mi5); memcpy(k.zd_der, de5, sizeof de5);
    }
    if (d.time_order == 4) {
        const double a[4] = {1., 1., 1., 1.}, b[4] = {0., 0., 0., 0.}, q[4] = {0.5, 0.5, 1., 1. / 6.};
        const double f[4] = {0., 0., 0., 1. / 6.}, m[4] = {1., 2., 2., 0.};
        memcpy(c->rk_a, a, sizeof a); memcpy(c->rk_b, b, sizeof b); memcpy(c->rk_q, q, sizeof q);
        memcpy(c->rk_f, f, sizeof f); memcpy(c->rk_m, m, sizeof m);
    } else {
        const double a[4] = {1., 3. / 4., 1. / 3., 0.}, b[4] = {0., 1. / 4., 2. / 3., 0.};
        const double q[4] = {1., 1. / 4., 2. / 3., 0.}, z[4] = {0., 0., 0., 0.};
        memcpy(c->rk_a, a, sizeof a); memcpy(c->rk_b, b, sizeof b); memcpy(c->rk_q, q, sizeof q);
        memcpy(c->rk_f, z, sizeof z); memcpy(c->rk_m, z, sizeof z);
    }

    hipError_t e = hipGetDevice(&c->dev);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) { delete c; return set_err(SHX_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    const size_t bytes = sizeof(double) * (size_t)gm.nalloc;
    std::vector<double **> all;
    for (int b = 0; b < 3; b++) for (int m = 0; m < 5; m++) all.push_back(&c->prim[b][m]);
    for (int m = 0; m < 5; m++) { all.push_back(&c->q[m]); all.push_back(&c->qsum[m]); }
    for (int m = 0; m < 10; m++) all.push_back(&c->met[m]);
    for (int m = 0; m < 12; m++) all.push_back(&c->grad[m]);
    all.push_back(&c->T);
    all.push_back(&c->mu);
    if (c->pc) { all.push_back(&c->P); all.push_back(&c->C); }
    if (c->rudy) { all.push_back(&c->PL); all.push_back(&c->RL); all.push_back(&c->UL); all.push_back(&c->C0); }
    for (double **p : all) {
        e = hipMalloc((void **)p, bytes);
        if (e == hipSuccess) e = hipMemsetAsync(*p, 0, bytes, c->stream);
        if (e != hipSuccess) {
            shx_destroy(c);
            return set_err(SHX_ERR_HIP, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        }
    }
    e = hipMalloc((void **)&c->d_red, sizeof(unsigned long long) * RED_SLOTS);
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_cnt, sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_nan, sizeof(int));
    if (e == hipSuccess) e = hipMemsetAsync(c->d_nan, 0, sizeof(int), c->stream);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_io, sizeof(shx_case::HostIo), hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { shx_destroy(c); return set_err(SHX_ERR_HIP, "setup: %s", hipGetErrorString(e)); }
    c->cur = 0;
    *out = c;
    return SHX_OK;
}

extern "C" int shx_destroy(shx_case *c)
{
    if (!c) return SHX_OK;
    if (c->batch) return set_err(SHX_ERR_STATE, "the case is in a batch: destroy the batch first");
    if (c->blocks) return set_err(SHX_ERR_STATE, "the case is a block of a multi-block case: destroy that first");
    if (c->stream) hipStreamSynchronize(c->stream);
    for (int b = 0; b < 3; b++) for (int m = 0; m < 5; m++) hipFree(c->prim[b][m]);
    for (int m = 0; m < 5; m++) { hipFree(c->q[m]); hipFree(c->qsum[m]); hipFree(c->src[m]); }
    for (int m = 0; m < 10; m++) hipFree(c->met[m]);
    for (int m = 0; m < 12; m++) hipFree(c->grad[m]);
    for (int m = 0; m < 6; m++) hipFree(c->bs[m]);
    hipFree(c->T); hipFree(c->mu); hipFree(c->d_red); hipFree(c->d_cnt); hipFree(c->d_nan);
    if (c->h_io) hipHostFree(c->h_io);
    for (int b = 0; b < 2; b++) {
        if (c->stage_ev[b]) hipEventDestroy(c->stage_ev[b]);
        if (c->stage_h[b]) hipHostFree(c->stage_h[b]);
    }
    hipFree(c->P); hipFree(c->C); hipFree(c->PL); hipFree(c->RL); hipFree(c->UL); hipFree(c->C0);
    for (int b = 0; b < 2; b++) {
        for (int m = 0; m < 5; m++) hipFree(c->bk[b][m]);
        hipFree(c->bkP[b]);
    }
    for (int m = 0; m < 3; m++) hipFree(c->post[m]);
    hipFree(c->film);
    hipFree(c->d_probes);
    hipFree(c->hist);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return SHX_OK;
}

extern "C" int shx_padded_extent(const shx_case *c, int ng[3])
{
    if (!c || !ng) return set_err(SHX_ERR_ARG, "null argument");
    for (int a = 0; a < 3; a++) ng[a] = c->gm.ng[a];
    return SHX_OK;
}

// host arrays use the reference layout (dense rows); device rows are padded to the pitch
static void row_geometry(const shx_case *c, size_t &width, size_t &height, size_t &dpitch)
{
    const Geom &gm = c->gm;
    if (gm.dims == 3) { width = sizeof(double) * gm.ng[2]; height = (size_t)gm.ng[0] * gm.ng[1]; dpitch = sizeof(double) * gm.st[1]; }
    else { width = sizeof(double) * gm.ng[1]; height = (size_t)gm.ng[0]; dpitch = sizeof(double) * gm.st[0]; }
}
// Host transfers go through two pinned staging buffers: a pitched copy between the padded device
// rows and dense rows in pinned memory (a DMA at PCIe rate), and a plain memcpy between pinned and
// the caller's memory, chunk by chunk, the memcpy of one chunk overlapping the DMA of the next.
// A pitched copy straight to or from pageable memory ran at 0.3 GB/s (256^3: 2.5 s to download the
// five state fields); staged, the transfer is bound by the host memcpy.
constexpr size_t STAGE_MAX = size_t(8) << 20;
// memcpy of a staging chunk on up to four host threads (one host thread copies ~14 GB/s, under the
// DMA's rate); chunks under 1 MB are copied on the calling thread
static void stage_memcpy(void *dst, const void *src, const size_t n)
{
    constexpr int NTH = 4;
    if (n < (size_t(1) << 20)) { memcpy(dst, src, n); return; }
    const size_t part = ((n + NTH - 1) / NTH + 63) & ~size_t(63);
    std::thread th[NTH - 1];
    int started = 0;
    for (int t = 1; t < NTH; t++) {
        const size_t off = (size_t)t * part;
        if (off >= n) break;
        char *d = (char *)dst + off;
        const char *sp = (const char *)src + off;
        const size_t len = std::min(part, n - off);
        try {
            th[started] = std::thread([d, sp, len] { memcpy(d, sp, len); });
            started++;
        } catch (...) {
            memcpy(d, sp, len);        // no thread available: copy this part here
        }
    }
    memcpy(dst, src, std::min(part, n));
    for (int t = 0; t < started; t++) th[t].join();
}
static int ensure_stage(shx_case *c, const size_t w, const size_t h)
{
    if (c->stage_h[0]) return SHX_OK;
    c->stage_bytes = std::max(w, std::min(STAGE_MAX, w * h));
    for (int b = 0; b < 2; b++) {
        HIPCHK(hipHostMalloc((void **)&c->stage_h[b], c->stage_bytes, hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&c->stage_ev[b], hipEventDisableTiming));
        HIPCHK(hipEventRecord(c->stage_ev[b], c->stream));
    }
    return SHX_OK;
}
static int copy_in(shx_case *c, double *dst, const double *src)
{
    if (!src) return set_err(SHX_ERR_ARG, "null field pointer");
    size_t w, h, dp;
    row_geometry(c, w, h, dp);
    int rc = ensure_stage(c, w, h);
    if (rc) return rc;
    const size_t rows = c->stage_bytes / w;
    int b = 0;
    for (size_t r0 = 0; r0 < h; r0 += rows, b ^= 1) {
        const size_t nr = std::min(rows, h - r0);
        // the buffer's previous copy (two chunks back, or an earlier call's) has been read
        HIPCHK(hipEventSynchronize(c->stage_ev[b]));
        stage_memcpy(c->stage_h[b], (const char *)src + r0 * w, nr * w);
        HIPCHK(hipMemcpy2DAsync((char *)dst + r0 * dp, dp, c->stage_h[b], w, w, nr, hipMemcpyHostToDevice,
                                c->stream));
        HIPCHK(hipEventRecord(c->stage_ev[b], c->stream));
    }
    return SHX_OK;
}
static int copy_out(shx_case *c, double *dst, const double *src)
{
    if (!dst) return set_err(SHX_ERR_ARG, "null field pointer");
    size_t w, h, dp;
    row_geometry(c, w, h, dp);
    int rc = ensure_stage(c, w, h);
    if (rc) return rc;
    const size_t rows = c->stage_bytes / w;
    // a staging buffer may still feed an upload's copy: the first chunk's DMA is queued behind it
    // on the stream, and every chunk is read out only after its own event
    size_t prev = h;     // first row of the chunk waiting to be read out (h: none)
    int b = 0;
    for (size_t r0 = 0;; r0 += rows, b ^= 1) {
        if (r0 < h) {
            const size_t nr = std::min(rows, h - r0);
            HIPCHK(hipMemcpy2DAsync(c->stage_h[b], w, (const char *)src + r0 * dp, dp, w, nr,
                                    hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipEventRecord(c->stage_ev[b], c->stream));
        }
        if (prev < h) {
            const int pb = b ^ 1;
            HIPCHK(hipEventSynchronize(c->stage_ev[pb]));
            stage_memcpy((char *)dst + prev * w, c->stage_h[pb], std::min(rows, h - prev) * w);
        }
        if (r0 >= h) break;
        prev = r0;
    }
    return SHX_OK;
}

static int ghost_fill(shx_case *c, double *const *f, int nf)
{
    Fields10 ff;
    memset(&ff, 0, sizeof ff);
    for (int m = 0; m < nf; m++) ff.f[m] = f[m];
    int rc = tbeg(c, KT_GHOST);
    if (rc) return rc;
    k_ghost_periodic<<<nblk(c->gm.ntot, 256), 256, 0, c->stream>>>(ff, nf, c->gm);
    LAUNCH_CHECK();
    return tend(c);
}

// face writers of every non-periodic face, in face order (i-, i+, j-, j+, k-, k+)
// argument record of the writer of face f (f = 2 * direction + upper) in a step of length tau;
// nodes: ghost nodes it owns
static void face_args(const shx_case *c, double *const *pr, int f, double tau, BcArgs &a, long &nodes)
{
    const Geom &gm = c->gm;
    const int kind = c->d.bc[f];
    memset(&a, 0, sizeof a);
    for (int m = 0; m < 5; m++) a.pr.f[m] = pr[m];
    if (!c->d.inviscid) { a.T = c->T; a.mu = c->mu; }
    for (int m = 0; m < 6; m++) a.bs[m] = c->bs[m];
    a.kind = kind;
    a.dir = f / 2;
    a.upper = f % 2;
    a.g = gm.g;
    a.t1 = (a.dir == 0) ? 1 : 0;
    a.t2 = (a.dir == 2) ? 1 : 2;
    if (kind == SHX_BC_MANUFACTURED || mirrors_metric(kind)) {
        // interior tangential range (src/ManufacturedSolution.c:202-338, src/BC.c:804-1172)
        a.lo1 = gm.lo[a.t1]; a.n1 = gm.n[a.t1]; a.lo2 = gm.lo[a.t2]; a.n2 = gm.n[a.t2];
    } else {
        a.lo1 = 0; a.n1 = gm.ng[a.t1]; a.lo2 = 0; a.n2 = gm.ng[a.t2];
    }
    a.first = gm.lo[a.dir];
    a.last = gm.lo[a.dir] + gm.n[a.dir] - 1;
    for (int d = 0; d < 3; d++) a.st[d] = gm.st[d];
    a.rho_in = c->d.rho_inflow; a.u_in = c->d.u_inflow; a.v_in = c->d.v_inflow;
    a.w_in = c->d.w_inflow; a.p_in = c->d.p_inflow;
    a.cinf = 1.0 / c->d.mach;
    a.mach = c->d.mach;
    a.T_wall = c->d.T_wall;
    for (int m = 0; m < 9; m++) a.met[m] = c->met[m];
    a.P = c->P; a.C = c->C;
    a.PL = c->PL; a.RL = c->RL; a.UL = c->UL; a.C0 = c->C0;
    a.p_out = c->d.p_out;
    a.alpha_tau = c->d.alpha_nonref * tau;
    a.k = c->k;
    nodes = (kind == SHX_BC_PERIODIC || kind == SHX_BC_INTERFACE) ? 0 : (long)a.n1 * a.n2 * a.g;
}
static int face_fill(shx_case *c, double *const *pr, double tau)
{
    if (!c->faces) return SHX_OK;
    int rc = tbeg(c, KT_FACES);
    if (rc) return rc;
    for (int f = 0; f < 2 * c->gm.dims; f++) {
        BcArgs a;
        long nodes;
        face_args(c, pr, f, tau, a, nodes);
        if (nodes == 0) continue;
        k_face_bc<<<nblk(nodes, 256), 256, 0, c->stream>>>(a);
        LAUNCH_CHECK();
    }
    return tend(c);
}
// start of a step on state pr: keep the ghost p, rho and u of every Rudy face
static int rudy_save(shx_case *c, double *const *pr)
{
    if (!c->rudy) return SHX_OK;
    for (int f = 0; f < 2 * c->gm.dims; f++) {
        if (c->d.bc[f] != SHX_BC_OUTFLOW_RUDY) continue;
        BcArgs a;
        long nodes;
        face_args(c, pr, f, 0.0, a, nodes);
        k_face_save<<<nblk(nodes, 256), 256, 0, c->stream>>>(a);
        LAUNCH_CHECK();
    }
    return SHX_OK;
}
static int carry_ghosts(shx_case *c, double *const *from, double *const *to)
{
    if (!c->faces) return SHX_OK;
    CFields5 a;
    Fields5 b;
    for (int m = 0; m < 5; m++) { a.f[m] = from[m]; b.f[m] = to[m]; }
    int rc = tbeg(c, KT_FACES);
    if (rc) return rc;
    k_carry_ghosts<<<nblk(c->gm.ntot, 256), 256, 0, c->stream>>>(a, b, c->gm);
    LAUNCH_CHECK();
    return tend(c);
}

extern "C" int shx_build_metric(shx_case *c, const double *xe, const double *ye, const double *ze)
{
    if (!c || !xe || !ye || (c->gm.dims == 3 && !ze)) return set_err(SHX_ERR_ARG, "null argument");
    const Geom &gm = c->gm;
    const long next = (long)(gm.ng[0] + 2) * (gm.ng[1] + 2) * (gm.dims == 3 ? gm.ng[2] + 2 : 1);
    const size_t bytes = sizeof(double) * (size_t)next;
    double *dx = nullptr, *dy = nullptr, *dz = nullptr;
    HIPCHK(hipMalloc((void **)&dx, bytes));
    HIPCHK(hipMalloc((void **)&dy, bytes));
    HIPCHK(hipMalloc((void **)&dz, bytes));
    HIPCHK(hipMemcpyAsync(dx, xe, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dy, ye, bytes, hipMemcpyHostToDevice, c->stream));
    if (gm.dims == 3) HIPCHK(hipMemcpyAsync(dz, ze, bytes, hipMemcpyHostToDevice, c->stream));
    else HIPCHK(hipMemsetAsync(dz, 0, bytes, c->stream));
    Fields10 mf;
    for (int m = 0; m < 10; m++) mf.f[m] = c->met[m];
    int rc = tbeg(c, KT_METRIC);
    if (rc) return rc;
    k_metric<<<nblk(gm.ntot, 256), 256, 0, c->stream>>>(dx, dy, dz, mf, gm,
                                                       (c->d.quirks & SHX_QUIRK_METRIC_DIAG3D) ? 1 : 0);
    LAUNCH_CHECK();
    if ((rc = tend(c))) return rc;
    if ((rc = ghost_fill(c, c->met, 10))) return rc;
    // wall and farfield faces, in face order: the ghost metric mirrored (full tangential range)
    for (int f = 0; f < 2 * gm.dims; f++) {
        if (!mirrors_metric(c->d.bc[f])) continue;
        BcArgs a;
        long nodes;
        face_args(c, c->q, f, 0.0, a, nodes);
        a.lo1 = 0; a.n1 = gm.ng[a.t1]; a.lo2 = 0; a.n2 = gm.ng[a.t2];
        nodes = (long)a.n1 * a.n2 * a.g;
        k_mirror_metric<<<nblk(nodes, 256), 256, 0, c->stream>>>(a, mf);
        LAUNCH_CHECK();
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    hipFree(dx); hipFree(dy); hipFree(dz);
    c->has_metric = 1;
    c->diag = (gm.dims == 3 && (c->d.quirks & SHX_QUIRK_METRIC_DIAG3D)) ? 1 : 0;
    return SHX_OK;
}

extern "C" int shx_upload_metric(shx_case *c, const double *const met[10])
{
    if (!c || !met) return set_err(SHX_ERR_ARG, "null argument");
    for (int m = 0; m < 10; m++) { int rc = copy_in(c, c->met[m], met[m]); if (rc) return rc; }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->has_metric = 1;
    // the diagonal-metric kernels apply when every off-diagonal entry is exactly zero
    int diag = (c->gm.dims == 3);
    static const int offd[6] = {1, 2, 3, 5, 6, 7};
    for (int s = 0; s < 6 && diag; s++)
        for (long n = 0; n < c->gm.ntot; n++)
            if (met[offd[s]][n] != 0.0) { diag = 0; break; }
    c->diag = diag;
    return SHX_OK;
}

extern "C" int shx_download_metric(shx_case *c, double *const met[10])
{
    if (!c || !met) return set_err(SHX_ERR_ARG, "null argument");
    for (int m = 0; m < 10; m++) { int rc = copy_out(c, met[m], c->met[m]); if (rc) return rc; }
    HIPCHK(hipStreamSynchronize(c->stream));
    return SHX_OK;
}

extern "C" int shx_upload_state(shx_case *c, const double *const prim[5])
{
    if (!c || !prim) return set_err(SHX_ERR_ARG, "null argument");
    // every stage buffer gets the full array: ghost nodes that no fill writes keep these values,
    // as they do in the reference's single U_RK array
    const size_t bytes = sizeof(double) * (size_t)c->gm.nalloc;
    for (int m = 0; m < 5; m++) {
        int rc = copy_in(c, c->prim[c->cur][m], prim[m]);
        if (rc) return rc;
        for (int b = 0; b < 3; b++)
            if (b != c->cur)
                HIPCHK(hipMemcpyAsync(c->prim[b][m], c->prim[c->cur][m], bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    if (c->pc) {
        // stored p and c of the uploaded state on every node; for Rudy faces also the step-start
        // sound speed, which stays as uploaded (the reference never recomputes U_lastStep's c)
        CFields5 cp;
        for (int m = 0; m < 5; m++) cp.f[m] = c->prim[c->cur][m];
        k_calc_values<<<nblk(c->gm.ntot, 256), 256, 0, c->stream>>>(cp, nullptr, nullptr, c->P, c->C, c->gm, c->k);
        LAUNCH_CHECK();
        if (c->rudy) HIPCHK(hipMemcpyAsync(c->C0, c->C, bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->has_state = 1;
    c->last_in = -1;
    return SHX_OK;
}

extern "C" int shx_upload_boundary_state(shx_case *c, const double *const bs[6])
{
    if (!c || !bs) return set_err(SHX_ERR_ARG, "null argument");
    const size_t bytes = sizeof(double) * (size_t)c->gm.nalloc;
    for (int m = 0; m < 6; m++) {
        if (!c->bs[m]) {
            HIPCHK(hipMalloc((void **)&c->bs[m], bytes));
            HIPCHK(hipMemsetAsync(c->bs[m], 0, bytes, c->stream));
        }
        int rc = copy_in(c, c->bs[m], bs[m]);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->has_bs = 1;
    return SHX_OK;
}

static int sync_ghosts(shx_case *c);
extern "C" int shx_download_state(shx_case *c, double *const prim[5])
{
    if (!c || !prim) return set_err(SHX_ERR_ARG, "null argument");
    // the ghost values the reference's U_lastStep holds after a step
    int rc0 = sync_ghosts(c);
    if (rc0) return rc0;
    for (int m = 0; m < 5; m++) { int rc = copy_out(c, prim[m], c->prim[c->cur][m]); if (rc) return rc; }
    HIPCHK(hipStreamSynchronize(c->stream));
    return SHX_OK;
}

extern "C" int shx_upload_source(shx_case *c, const double *const src[5])
{
    if (!c) return set_err(SHX_ERR_ARG, "null argument");
    if (!src) { c->has_src = 0; return SHX_OK; }
    const size_t bytes = sizeof(double) * (size_t)c->gm.nalloc;
    for (int m = 0; m < 5; m++) {
        if (!c->src[m]) HIPCHK(hipMalloc((void **)&c->src[m], bytes));
        int rc = copy_in(c, c->src[m], src[m]);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->has_src = 1;
    return SHX_OK;
}

// A kernel choice for one launch class: the single-case kernel, its batched twin and the tile.
struct Pick {
    const void *single, *batch, *batchp;   // single case, batched flat, batched persistent
    int TA, TDI;
};

template <int S, int TA, int TDI, int AX, bool MIX>
static Pick weno_pick_t()
{
    return {(const void *)&k_weno<S, TA, TDI, AX, MIX>, (const void *)&k_weno_batch<S, TA, TDI, AX, MIX, false>,
            (const void *)&k_weno_batch<S, TA, TDI, AX, MIX, true>, TA, TDI};
}
// tile variants of the WENO sweeps (TA transverse x TDI along the sweep), chosen per case and
// direction by shx_tune: sweeps along the contiguous axis 4 x 64 or 2 x 128, the others 16 x 16
// or 8 x 32.  Diagonal metrics only ever use AX = 2 on the contiguous axis and 0 or 1 otherwise.
constexpr int NWTILE = 2;
template <int S, int TA, int TDI, bool MIX, bool COAL>
static Pick weno_pick_m(int ax)
{
    if constexpr (COAL) {
        if (ax == 2) return weno_pick_t<S, TA, TDI, 2, MIX>();
        return weno_pick_t<S, TA, TDI, -1, MIX>();
    } else {
        if (ax == 0) return weno_pick_t<S, TA, TDI, 0, MIX>();
        if (ax == 1) return weno_pick_t<S, TA, TDI, 1, MIX>();
        return weno_pick_t<S, TA, TDI, -1, MIX>();
    }
}
template <int S, int TA, int TDI, bool COAL>
static Pick weno_pick_s(int ax, bool mix)
{
    return mix ? weno_pick_m<S, TA, TDI, true, COAL>(ax) : weno_pick_m<S, TA, TDI, false, COAL>(ax);
}
template <int S>
static Pick weno_pick_v(bool coalesce, int variant, int ax, bool mix)
{
    if (coalesce) return variant ? weno_pick_s<S, 2, 128, true>(ax, mix) : weno_pick_s<S, 4, 64, true>(ax, mix);
    return variant ? weno_pick_s<S, 8, 32, false>(ax, mix) : weno_pick_s<S, 16, 16, false>(ax, mix);
}
static Pick weno_pick(const shx_case *c, int dir)
{
    const bool diag = (c->diag && c->gm.dims == 3);
    const bool mix = (c->d.precision == SHX_MIXED);
    const bool coalesce = (c->gm.st[dir] == 1);
    const int ax = diag ? dir : -1;
    const int v = c->wtile[dir];
    return (c->d.space_order == 5) ? weno_pick_v<5>(coalesce, v, ax, mix) : weno_pick_v<9>(coalesce, v, ax, mix);
}

// transverse axes of a sweep: A = the contiguous non-sweep axis with extent > 1, B = the other
static void sweep_axes(const Geom &gm, int dir, int &A, int &B)
{
    if (gm.dims == 3) {
        if (dir == 2) { A = 1; B = 0; } else { A = 2; B = (dir == 0) ? 1 : 0; }
    } else {
        A = (dir == 0) ? 1 : 0; B = 2;
    }
}

static WenoArgs weno_args(const shx_case *c, double *const *pr, int dir, int first)
{
    const Geom &gm = c->gm;
    WenoArgs a;
    memset(&a, 0, sizeof a);
    for (int m = 0; m < 5; m++) { a.pr.f[m] = pr[m]; a.q.f[m] = c->q[m]; }
    const bool diag = (c->diag && gm.dims == 3);
    if (diag) { a.m0 = c->met[4 * dir]; a.m1 = nullptr; a.m2 = nullptr; }
    else { a.m0 = c->met[3 * dir]; a.m1 = c->met[3 * dir + 1]; a.m2 = c->met[3 * dir + 2]; }
    a.J = c->met[9];
    int A, B;
    sweep_axes(gm, dir, A, B);
    a.sD = gm.st[dir]; a.sA = gm.st[A]; a.sB = gm.st[B];
    a.nD = gm.n[dir]; a.nA = gm.n[A]; a.nB = gm.n[B];
    a.oD = gm.lo[dir]; a.oA = gm.lo[A]; a.oB = gm.lo[B];
    a.ngD = gm.ng[dir];
    a.first = first;
    a.coalesce_d = (a.sD == 1);
    a.alignA = (a.sA == 1);
    a.P = c->P; a.C = c->C;
    a.pc_lo = stores_pc(c->d.bc[2 * dir]);
    a.pc_hi = stores_pc(c->d.bc[2 * dir + 1]);
    a.w_lo = is_wall(c->d.bc[2 * dir]);
    a.w_hi = is_wall(c->d.bc[2 * dir + 1]);
    wall_corrector(c->d.bc[2 * dir], dir, a.c_lo);
    wall_corrector(c->d.bc[2 * dir + 1], dir, a.c_hi);
    a.k = c->k;
    return a;
}

// tile grid of a sweep: tiles along A cover the padded span when A is contiguous (aligned rows)
static dim3 sweep_grid(int nA, int oA, int alignA, int nB, int nD, const Pick &p)
{
    const int spanA = alignA ? oA + nA : nA;
    return dim3(nblk(spanA, p.TA), nB, nblk(nD, p.TDI - 1));
}

static int launch_pick(const void *fn, dim3 grid, int threads, void *arg, hipStream_t st)
{
    void *args[] = {arg};
    HIPCHK(hipLaunchKernel(fn, grid, dim3(threads), args, 0, st));
    return SHX_OK;
}

// one WENO sweep along axis dir of the state `pr`
// one-dimensional launch of a tile grid g; tg receives its extents and the tile order
static dim3 flat_grid(const dim3 g, int tg[4], int order)
{
    tg[0] = (int)g.x; tg[1] = (int)g.y; tg[2] = (int)g.z; tg[3] = order;
    return dim3(g.x * g.y * g.z);
}
// tile order of a sweep (see sweep_tile): along the sweep, except the viscous sweep of the
// slowest axis
static int sweep_order(bool visc, int dir) { return (visc && dir == 0) ? 0 : 1; }
static int weno_sweep(shx_case *c, double *const *pr, int dir, int first)
{
    WenoArgs a = weno_args(c, pr, dir, first);
    const Pick p = weno_pick(c, dir);
    int rc = tbeg(c, KT_WENO0 + dir);
    if (rc) return rc;
    const dim3 grid = flat_grid(sweep_grid(a.nA, a.oA, a.alignA, a.nB, a.nD, p), a.tg, sweep_order(false, dir));
    if ((rc = launch_pick(p.single, grid, p.TA * p.TDI, &a, c->stream))) return rc;
    return tend(c);
}

// residual of one stage into c->q: ghost fill, derived quantities, viscous derivatives, WENO
// sweeps of every direction, then the viscous sweeps (the reference's accumulation order,
// src/Functions.c:2161-2272)
static int visc_sweep(shx_case *c, double *const *pr, int dir, const RkArgs *rk);
static int zeta_sweep(shx_case *c, double *const *pr, const RkArgs *rk);
// the zeta direction runs as one fused kernel (k_zeta)
static bool fuse_zeta(const shx_case *c)
{
    return c->gm.dims == 3 && c->diag && !c->d.inviscid && c->wtile[2] == 0;
}
// the viscous xi direction rides on the derivative march (k_visc_grad_march_xi; WENO5 order, whose
// rings fit the march's registers)
static bool fuse_xi(const shx_case *c)
{
    return c->gm.dims == 3 && c->diag && !c->d.inviscid && c->d.space_order == 5;
}

// the eta direction runs as one fused kernel (k_eta) where the viscous xi terms ride on the march
// and the zeta direction is fused too
static bool fuse_eta(const shx_case *c) { return fuse_xi(c) && fuse_zeta(c); }

static double iso_factor(const shx_case *c);
// argument record of the 3D derivative march on stage input pr; xi: with the viscous xi terms
static GradArgs grad_args(const shx_case *c, double *const *pr, const bool xi)
{
    GradArgs ga;
    memset(&ga, 0, sizeof ga);
    for (int m = 0; m < 5; m++) ga.pr.f[m] = pr[m];
    for (int m = 0; m < 12; m++) ga.gr.f[m] = c->grad[m];
    ga.T = c->T; ga.gm = c->gm; ga.k = c->k; ga.ci = 32; ga.tgrad = c->diag ? 0 : 1;
    if (xi) {
        for (int m = 0; m < 5; m++) ga.q.f[m] = c->q[m];
        ga.mu = c->mu; ga.m0 = c->met[0]; ga.m4 = c->met[4]; ga.m8 = c->met[8]; ga.J = c->met[9];
        ga.iso_hi = (c->d.bc[1] == SHX_BC_WALL_ISOTHERMAL) ? 1 : 0;
        ga.gam_iso = iso_factor(c);
    }
    return ga;
}
static dim3 march_grid(const Geom &gm, const bool xi)
{
    return dim3(nblk(gm.ng[2], xi ? GKX : GK), nblk(gm.ng[1], xi ? GJX : GJ), nblk(gm.ng[0], 32));
}
// Nodal derivatives of the viscous terms (3D: the march; 2D: pointwise).  xi: the march also adds
// the viscous xi terms to Q (fuse_xi cases, after the WENO xi sweep).
static int grad_march(shx_case *c, double *const *pr, const bool xi)
{
    const Geom &gm = c->gm;
    int rc = tbeg(c, xi ? KT_GRADXI : KT_GRAD);
    if (rc) return rc;
    if (gm.dims == 3) {
        GradArgs ga = grad_args(c, pr, xi);
        const void *fn = xi ? (const void *)&k_visc_grad_march_xi<5> : march_kernel(c->d.space_order, ga.tgrad, false);
        const dim3 grid = flat_grid(march_grid(gm, xi), ga.tg, 0);
        void *args[] = {&ga};
        HIPCHK(hipLaunchKernel(fn, grid, dim3(xi ? GKX * GJX : GK * GJ), args, 0, c->stream));
    } else {
        CFields5 cp;
        Fields12 gf;
        for (int m = 0; m < 5; m++) cp.f[m] = pr[m];
        for (int m = 0; m < 12; m++) gf.f[m] = c->grad[m];
        k_visc_grad<<<nblk(gm.ntot, 256), 256, 0, c->stream>>>(cp, c->T, gf, gm, c->k);
    }
    LAUNCH_CHECK();
    return tend(c);
}
static int eta_sweep(shx_case *c, double *const *pr);
// rk: when given, the last viscous sweep applies this stage update instead of storing Q
static int stage_fluxes(shx_case *c, double *const *pr, const RkArgs *rk, double tau)
{
    const Geom &gm = c->gm;
    int rc;
    if (c->d.inviscid && !c->pc) {
        if ((rc = ghost_fill(c, pr, 5))) return rc;
    } else {
        // periodic ghosts, then T and mu (viscous) and the stored p and c (subsonic outflow faces)
        Fields5 fp;
        for (int m = 0; m < 5; m++) fp.f[m] = pr[m];
        const bool visc = !c->d.inviscid;
        if ((rc = tbeg(c, KT_VALUES))) return rc;
        k_ghost_values<<<nblk(gm.ntot, 256), 256, 0, c->stream>>>(fp, visc ? c->T : nullptr, visc ? c->mu : nullptr,
                                                                    c->P, c->C, gm, c->k);
        LAUNCH_CHECK();
        if ((rc = tend(c))) return rc;
    }
    const bool fx = fuse_xi(c);
    if (!c->d.inviscid) {
        if ((rc = face_fill(c, pr, tau))) return rc;
        if (!fx && (rc = grad_march(c, pr, false))) return rc;
    } else if ((rc = face_fill(c, pr, tau))) {
        return rc;
    }
    // Q accumulates the WENO divergences, then the viscous terms; with the fused zeta direction
    // the order is WENO xi, eta, viscous xi, eta, then WENO zeta and viscous zeta together.  With
    // the viscous xi direction in the derivative march, the march runs after the WENO xi and eta
    // sweeps, in the place of the xi viscous sweep.
    // With the eta direction fused as well, the order is WENO xi, the march with viscous xi, then
    // the fused eta and zeta kernels (each WENO and viscous together).
    const bool fz = fuse_zeta(c), fe = fuse_eta(c);
    const int nsep = fz ? gm.dims - 1 : gm.dims;
    for (int dir = 0; dir < nsep; dir++) {
        if (fe && dir == 1) continue;
        if ((rc = weno_sweep(c, pr, dir, dir == 0))) return rc;
    }
    if (fx && (rc = grad_march(c, pr, true))) return rc;
    if (!c->d.inviscid)
        for (int dir = fx ? 1 : 0; dir < nsep; dir++) {
            if (fe && dir == 1) rc = eta_sweep(c, pr);
            else rc = visc_sweep(c, pr, dir, dir == gm.dims - 1 ? rk : nullptr);
            if (rc) return rc;
        }
    if (fz && (rc = zeta_sweep(c, pr, rk))) return rc;
    return SHX_OK;
}

template <int S, int DIR, int TA, int TDI, bool DIAG>
static Pick visc_pick_t()
{
    return {(const void *)&k_visc_sweep<S, DIR, TA, TDI, DIAG>, (const void *)&k_visc_batch<S, DIR, TA, TDI, DIAG, false>,
            (const void *)&k_visc_batch<S, DIR, TA, TDI, DIAG, true>, TA, TDI};
}

// tiles sized so that the LDS window (13 or 23 values per node) plus the face buffer fit:
// diagonal metric 16 x 16 for WENO5 and 16 x 32 for WENO9, whose 2G = 10 halo rows make short
// tiles re-read most of their window (16 x 32 runs the WENO9 xi / eta sweeps 21 % faster than
// 16 x 12 at 384^3; WENO5 is fastest at 16 x 16); generic 16 x 12 / 16 x 8 (4 x 48 contiguous)
template <int S>
static Pick visc_pick_s(int dir, bool coalesce, bool diag)
{
    constexpr int TDN = (S == 5) ? 16 : 32;
    constexpr int TDG = (S == 5) ? 12 : 8;
    if (coalesce) {
        if (dir == 1) return visc_pick_t<S, 1, 4, 48, false>();       // 2D eta sweep
        if (diag) return visc_pick_t<S, 2, 4, 64, true>();
        return visc_pick_t<S, 2, 4, 48, false>();
    }
    if (diag) return (dir == 0) ? visc_pick_t<S, 0, 16, TDN, true>() : visc_pick_t<S, 1, 16, TDN, true>();
    return (dir == 0) ? visc_pick_t<S, 0, 16, TDG, false>() : visc_pick_t<S, 1, 16, TDG, false>();
}

static Pick visc_pick(const shx_case *c, int dir)
{
    const bool diag = c->diag && c->gm.dims == 3;
    const bool coalesce = (c->gm.st[dir] == 1);
    return (c->d.space_order == 5) ? visc_pick_s<5>(dir, coalesce, diag) : visc_pick_s<9>(dir, coalesce, diag);
}

// heat-conduction factor the isothermal writer gives its ghosts (Prandtl 0.72,
// src/BC.c:2974-2976), averaged with the interior one as src/ZD.c:193-194
static double iso_factor(const shx_case *c)
{
    const shx_desc &d = c->d;
    const double gw = 1.0 / ((d.gamma - 1.0) * pow(d.mach, 2.0) * d.reynolds * 0.72);
    return 0.5 * (gw + c->k.gam_heat);
}

static ViscArgs visc_args(const shx_case *c, double *const *pr, int dir, const RkArgs *rk)
{
    const Geom &gm = c->gm;
    ViscArgs a;
    memset(&a, 0, sizeof a);
    a.u = pr[1]; a.v = pr[2]; a.w = pr[3]; a.T = c->T; a.mu = c->mu;
    const int b1 = (dir == 0) ? 1 : 0, b2 = (dir == 2) ? 1 : 2;
    const bool diag = c->diag && gm.dims == 3;
    if (diag) {
        a.cr[0] = c->grad[b1 * 3 + b1]; a.cr[1] = c->grad[dir * 3 + b1];
        a.cr[2] = c->grad[b2 * 3 + b2]; a.cr[3] = c->grad[dir * 3 + b2];
    } else {
        for (int q = 0; q < 3; q++) { a.cr[q] = c->grad[q * 3 + b1]; a.cr[4 + q] = c->grad[q * 3 + b2]; }
        a.cr[3] = c->grad[9 + b1]; a.cr[7] = c->grad[9 + b2];
    }
    for (int m = 0; m < 10; m++) a.met.f[m] = c->met[m];
    for (int m = 0; m < 5; m++) a.q.f[m] = c->q[m];
    int A, B;
    sweep_axes(gm, dir, A, B);
    a.sD = gm.st[dir]; a.sA = gm.st[A]; a.sB = gm.st[B];
    a.nD = gm.n[dir]; a.nA = gm.n[A]; a.nB = gm.n[B];
    a.oD = gm.lo[dir]; a.oA = gm.lo[A]; a.oB = gm.lo[B];
    a.ngD = gm.ng[dir];
    a.ngA = gm.ng[A];
    a.coalesce_d = (a.sD == 1);
    a.alignA = (a.sA == 1);
    a.three = (gm.dims == 3);
    a.bug = (dir == 2 && (c->d.quirks & SHX_QUIRK_ZETA_VISC_J)) ? 1 : 0;
    a.k = c->k;
    a.iso_hi = (c->d.bc[2 * dir + 1] == SHX_BC_WALL_ISOTHERMAL) ? 1 : 0;
    a.gam_iso = iso_factor(c);
    a.fuse_rk = rk ? 1 : 0;
    if (rk) a.rk = *rk;
    return a;
}

// viscous face sweep along axis dir: Q += viscous flux divergence (after the WENO sweeps)
static int visc_sweep(shx_case *c, double *const *pr, int dir, const RkArgs *rk)
{
    ViscArgs a = visc_args(c, pr, dir, rk);
    const Pick p = visc_pick(c, dir);
    int rc = tbeg(c, KT_VISC0 + dir);
    if (rc) return rc;
    const dim3 grid = flat_grid(sweep_grid(a.nA, a.oA, a.alignA, a.nB, a.nD, p), a.tg, sweep_order(true, dir));
    if ((rc = launch_pick(p.single, grid, p.TA * p.TDI, &a, c->stream))) return rc;
    return tend(c);
}
template <int S, bool MIX>
static Pick zeta_pick_t()
{
    return {(const void *)&k_zeta<S, MIX>, (const void *)&k_zeta_batch<S, MIX, false>,
            (const void *)&k_zeta_batch<S, MIX, true>, 4, 64};
}
static Pick zeta_pick(const shx_case *c)
{
    const bool mix = (c->d.precision == SHX_MIXED);
    if (c->d.space_order == 5) return mix ? zeta_pick_t<5, true>() : zeta_pick_t<5, false>();
    return mix ? zeta_pick_t<9, true>() : zeta_pick_t<9, false>();
}
static ZetaArgs zeta_args(const shx_case *c, double *const *pr, const RkArgs *rk)
{
    ZetaArgs z;
    z.w = weno_args(c, pr, 2, 0);
    z.v = visc_args(c, pr, 2, rk);
    return z;
}
static int zeta_sweep(shx_case *c, double *const *pr, const RkArgs *rk)
{
    ZetaArgs z = zeta_args(c, pr, rk);
    const Pick p = zeta_pick(c);
    int rc = tbeg(c, KT_ZETA);
    if (rc) return rc;
    const dim3 grid = flat_grid(sweep_grid(z.w.nA, z.w.oA, z.w.alignA, z.w.nB, z.w.nD, p), z.w.tg, 1);
    if ((rc = launch_pick(p.single, grid, p.TA * p.TDI, &z, c->stream))) return rc;
    return tend(c);
}

template <bool MIX>
static Pick eta_pick_t()
{
    return {(const void *)&k_eta<MIX>, (const void *)&k_eta_batch<MIX, false>, (const void *)&k_eta_batch<MIX, true>, 8, 32};
}
static Pick eta_pick(const shx_case *c)
{
    return (c->d.precision == SHX_MIXED) ? eta_pick_t<true>() : eta_pick_t<false>();
}
static EtaArgs eta_args(const shx_case *c, double *const *pr)
{
    EtaArgs z;
    z.w = weno_args(c, pr, 1, 0);
    z.v = visc_args(c, pr, 1, nullptr);
    return z;
}
static int eta_sweep(shx_case *c, double *const *pr)
{
    EtaArgs z = eta_args(c, pr);
    const Pick p = eta_pick(c);
    int rc = tbeg(c, KT_ETA);
    if (rc) return rc;
    const dim3 grid = flat_grid(sweep_grid(z.w.nA, z.w.oA, z.w.alignA, z.w.nB, z.w.nD, p), z.w.tg, 1);
    if ((rc = launch_pick(p.single, grid, p.TA * p.TDI, &z, c->stream))) return rc;
    return tend(c);
}

static RkArgs rk_args(shx_case *c, double *const *us, double *const *out, int s, double tau)
{
    RkArgs A;
    memset(&A, 0, sizeof A);
    for (int m = 0; m < 5; m++) {
        A.un.f[m] = c->prim[c->cur][m]; A.us.f[m] = us[m]; A.out.f[m] = out[m];
        A.qsum.f[m] = c->qsum[m]; A.src.f[m] = c->src[m];
    }
    A.a = c->rk_a[s]; A.b = c->rk_b[s]; A.qf = c->rk_q[s]; A.ff = c->rk_f[s]; A.mf = c->rk_m[s];
    A.tau = tau;
    A.first_stage = (s == 0);
    A.has_src = c->has_src;
    A.three = (c->gm.dims == 3);
    A.gamma = c->k.gamma;
    A.upsilon = c->k.upsilon;
    return A;
}

static int rk_update(shx_case *c, const RkArgs &rk)
{
    const Geom &gm = c->gm;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    UpdArgs A;
    memset(&A, 0, sizeof A);
    for (int m = 0; m < 5; m++) A.q.f[m] = c->q[m];
    for (int m = 0; m < 10; m++) A.met.f[m] = c->met[m];
    A.J = c->met[9];
    A.gm = gm;
    A.rk = rk;
    int rc = tbeg(c, KT_RK);
    if (rc) return rc;
    k_rk_update<<<nblk(nint, 256), 256, 0, c->stream>>>(A);
    LAUNCH_CHECK();
    return tend(c);
}

static int check_ready(shx_case *c)
{
    if (!c) return set_err(SHX_ERR_ARG, "null handle");
    if (!c->has_metric) return set_err(SHX_ERR_STATE, "no metric: call shx_build_metric or shx_upload_metric");
    if (!c->has_state) return set_err(SHX_ERR_STATE, "no state: call shx_upload_state");
    if (c->needs_bs && !c->has_bs)
        return set_err(SHX_ERR_STATE, "manufactured faces need shx_upload_boundary_state");
    return SHX_OK;
}

// end-of-step checks fused into the last stage's update (step_rk flags)
enum { STEP_NAN = 1, STEP_CFL = 2 };

// One time step (CalcRungeKutta and the step-end copy).  flags: STEP_NAN adds checkNAN on the new
// state to the running count c->d_nan; STEP_CFL puts the maximum node CFL sum of the new state
// into c->d_red (which the caller clears).
// A step in three parts, so that the blocks of a multi-block case advance stage by stage together:
// step_begin (the Rudy faces' step-start state; buffers), step_stage (one stage), step_end (the
// step-end copy).  Stage s reads buffer st_in and writes st_out; stage s+1 reads what stage s wrote
// and writes the other scratch buffer.
static int step_begin(shx_case *c)
{
    const int un = c->cur;
    c->st_bx = (un + 1) % 3;
    c->st_by = (un + 2) % 3;
    c->st_in = un;
    c->st_out = c->st_bx;
    return rudy_save(c, c->prim[un]);
}
static int step_stage(shx_case *c, int s, double tau, int flags)
{
    int rc;
    RkArgs rk = rk_args(c, c->prim[c->st_in], c->prim[c->st_out], s, tau);
    if (s == c->d.time_order - 1) {
        if (flags & STEP_NAN) { rk.nan_cnt = c->d_nan; rk.nanP = c->P; }
        if (flags & STEP_CFL) rk.cfl_red = c->d_red;
    }
    // viscous stages apply the update in their last sweep; inviscid ones in k_rk_update
    if ((rc = stage_fluxes(c, c->prim[c->st_in], c->d.inviscid ? nullptr : &rk, tau))) return rc;
    if ((rc = carry_ghosts(c, c->prim[c->st_in], c->prim[c->st_out]))) return rc;
    if (c->d.inviscid && (rc = rk_update(c, rk))) return rc;
    c->st_in = c->st_out;
    c->st_out = (c->st_out == c->st_bx) ? c->st_by : c->st_bx;
    return SHX_OK;
}
static void step_end(shx_case *c)
{
    // the last stage's input holds the ghost values the reference's U_RK (and so U_lastStep) keeps
    // after the step; post-processing copies them over when it needs them
    c->last_in = (c->st_in == c->st_bx) ? c->st_by : c->st_bx;
    c->cur = c->st_in;   // U_lastStep <- U_RK (pointer swap instead of WriteValuesFromU1ToU2)
}
static int step_rk(shx_case *c, double tau, int flags)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (c->batch) return set_err(SHX_ERR_STATE, "the case is in a batch: step it with shx_batch_run");
    if (c->iface) return set_err(SHX_ERR_STATE, "the case has interface faces: step it with shx_blocks_run");
    if ((rc = step_begin(c))) return rc;
    for (int s = 0; s < c->d.time_order; s++)
        if ((rc = step_stage(c, s, tau, flags))) return rc;
    step_end(c);
    return SHX_OK;
}

extern "C" int shx_calc_runge_kutta(shx_case *c, double tau) { return step_rk(c, tau, 0); }

// Non-finite densities the fused checkNAN found since the last call (synchronizes the stream).
static int nan_found(shx_case *c, int &fresh)
{
    HIPCHK(hipMemcpyAsync(&c->h_io->nan, c->d_nan, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    fresh = c->h_io->nan - c->nan_seen;
    c->nan_seen = c->h_io->nan;
    return SHX_OK;
}

// ---- multi-block cases on one device -----------------------------------------------------------
// The blocks of a case share the GPU; before every stage one gather fills the ghost layers of every
// interface face of every block from the donor blocks' stage input (the reference's
// TransferFlowParameterWithGhosts, src/Functions.c:3580-3684, whose repeated ghost-inclusive passes
// settle edges and corners).  A ghost node is followed through the faces it lies beyond, one at a
// time in axis order, periodic and interface faces first: a periodic face wraps the index, an
// interface face maps it into the donor block (normal layer l outside the face = layer l inside the
// donor's face; tangential indices through the transform).  The value is read where no such face
// is left: an interior node, or a ghost node beyond boundary-writer faces only, whose value is the
// donor's writer output of the previous stage (carried into the stage buffer), as in the reference,
// whose transfers copy the donors' ghost-inclusive planes before its writers run.  A node that is
// read is never one that is written (written nodes lie beyond an interface face), so one launch
// serves all blocks.
enum { XK_WRITER = 0, XK_PERIODIC = 1, XK_INTERFACE = 2 };
struct XRec {
    Geom gm;
    gdp f[10];            // fields moved: the stage input rho,u,v,w,e, or the metric rows and J
    int kind[6], donor[6], dface[6], tr[6][3];
    long first;           // first work item of this block
};
struct XTab {
    XRec r[SHX_BLOCKS_MAX];
    int nb, metric;
    long total;
};
__global__ __launch_bounds__(256) void k_interface_gather(XTab t)
{
    const long w = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= t.total) return;
    int b0 = 0;
    while (b0 + 1 < t.nb && w >= t.r[b0 + 1].first) b0++;
    const XRec &R0 = t.r[b0];
    const int dims = R0.gm.dims;
    int x[3];
    const long n0 = padded_node(R0.gm, w - R0.first, x);
    bool across = false;                 // a ghost in the direction of an interface face
    for (int d = 0; d < dims; d++) {
        const int lo = R0.gm.lo[d], hi = lo + R0.gm.n[d];
        if (x[d] < lo && R0.kind[2 * d] == XK_INTERFACE) across = true;
        if (x[d] >= hi && R0.kind[2 * d + 1] == XK_INTERFACE) across = true;
    }
    if (!across) return;
    int b = b0;
    int perm[3] = {0, 1, 2};             // axis of block b that each axis of block b0 runs along
    double sgn[3] = {1.0, 1.0, 1.0};
    bool moved = false;
    for (int it = 0; it < 2 * dims; it++) {
        const XRec &R = t.r[b];
        int a = -1, up = 0;
        for (int d = 0; d < dims && a < 0; d++) {
            int u = -1;
            if (x[d] < R.gm.lo[d]) u = 0;
            else if (x[d] >= R.gm.lo[d] + R.gm.n[d]) u = 1;
            if (u >= 0 && R.kind[2 * d + u] != XK_WRITER) { a = d; up = u; }
        }
        if (a < 0) break;
        moved = true;
        const int f = 2 * a + up;
        if (R.kind[f] == XK_PERIODIC) { x[a] += up ? -R.gm.n[a] : R.gm.n[a]; continue; }
        const XRec &D = t.r[R.donor[f]];
        const int fd = R.dface[f], ad = fd >> 1, dup = fd & 1;
        const int l = up ? x[a] - (R.gm.lo[a] + R.gm.n[a]) : R.gm.lo[a] - 1 - x[a];
        int y[3] = {0, 0, 0}, m[3] = {0, 1, 2};
        double s[3] = {1.0, 1.0, 1.0};
        y[ad] = dup ? D.gm.lo[ad] + D.gm.n[ad] - 1 - l : D.gm.lo[ad] + l;
        m[a] = ad;
        s[a] = (up != dup) ? 1.0 : -1.0;
        for (int q = 0; q < dims; q++) {
            if (q == a) continue;
            const int tq = R.tr[f][q], aq = (tq > 0 ? tq : -tq) - 1, r = x[q] - R.gm.lo[q];
            y[aq] = (tq > 0) ? D.gm.lo[aq] + r : D.gm.lo[aq] + D.gm.n[aq] - 1 - r;
            m[q] = aq;
            s[q] = (tq > 0) ? 1.0 : -1.0;
        }
        for (int q = 0; q < dims; q++) { sgn[q] *= s[perm[q]]; perm[q] = m[perm[q]]; }
        for (int q = 0; q < 3; q++) x[q] = y[q];
        b = R.donor[f];
    }
    if (!moved) return;
    const XRec &S = t.r[b];
    for (int d = 0; d < dims; d++)
        if (x[d] < 0 || x[d] >= S.gm.ng[d]) return;
    const long ns = (long)x[0] * S.gm.st[0] + (long)x[1] * S.gm.st[1] + (long)x[2] * S.gm.st[2];
    if (!t.metric) {
#pragma unroll
        for (int m = 0; m < 5; m++) R0.f[m][n0] = S.f[m][ns];
    } else {
        // row a of the block's metric is the donor's row perm[a], signed (src/Functions.c:2664-2686)
        for (int a = 0; a < 3; a++) {
            const int pa = (a < dims) ? perm[a] : a;
            const double sa = (a < dims) ? sgn[a] : 1.0;
#pragma unroll
            for (int c = 0; c < 3; c++) R0.f[3 * a + c][n0] = sa * S.f[3 * pa + c][ns];
        }
        R0.f[9][n0] = S.f[9][ns];
    }
}

struct shx_blocks {
    std::vector<shx_case *> b;
    int kind[SHX_BLOCKS_MAX][6], donor[SHX_BLOCKS_MAX][6], dface[SHX_BLOCKS_MAX][6], tr[SHX_BLOCKS_MAX][6][3];
    hipStream_t stream = nullptr;
    hipEvent_t ev_x = nullptr;
    std::vector<hipEvent_t> ev_b;
};

// the gather's table: the stage input buffers of every block (metric: its metric arrays)
static XTab xtab(const shx_blocks *g, bool metric)
{
    XTab t;
    memset(&t, 0, sizeof t);
    t.nb = (int)g->b.size();
    t.metric = metric ? 1 : 0;
    long first = 0;
    for (int i = 0; i < t.nb; i++) {
        const shx_case *c = g->b[i];
        XRec &r = t.r[i];
        r.gm = c->gm;
        for (int m = 0; m < 10; m++) r.f[m] = metric ? c->met[m] : (m < 5 ? c->prim[c->st_in][m] : nullptr);
        for (int f = 0; f < 6; f++) {
            r.kind[f] = g->kind[i][f]; r.donor[f] = g->donor[i][f]; r.dface[f] = g->dface[i][f];
            for (int q = 0; q < 3; q++) r.tr[f][q] = g->tr[i][f][q];
        }
        r.first = first;
        first += c->gm.ntot;
    }
    t.total = first;
    return t;
}
static int gather(shx_blocks *g, bool metric)
{
    XTab t = xtab(g, metric);
    k_interface_gather<<<nblk(t.total, 256), 256, 0, g->stream>>>(t);
    LAUNCH_CHECK();
    return SHX_OK;
}

extern "C" int shx_blocks_create(shx_case *const *blocks, int nblocks, const shx_interface *ifc, int nifc,
                                 shx_blocks **out)
{
    if (!blocks || !out || nblocks <= 0 || (nifc > 0 && !ifc) || nifc < 0) return set_err(SHX_ERR_ARG, "null argument or no blocks");
    *out = nullptr;
    if (nblocks > SHX_BLOCKS_MAX) return set_err(SHX_ERR_UNSUPPORTED, "at most %d blocks", SHX_BLOCKS_MAX);
    const shx_case *c0 = blocks[0];
    for (int i = 0; i < nblocks; i++) {
        const shx_case *c = blocks[i];
        if (!c) return set_err(SHX_ERR_ARG, "block %d is null", i);
        for (int j = 0; j < i; j++)
            if (blocks[j] == c) return set_err(SHX_ERR_ARG, "block %d appears twice", i);
        if (c->blocks || c->batch) return set_err(SHX_ERR_STATE, "block %d belongs to another group or a batch", i);
        if (c->gm.dims != c0->gm.dims || c->d.space_order != c0->d.space_order || c->d.time_order != c0->d.time_order ||
            c->d.inviscid != c0->d.inviscid)
            return set_err(SHX_ERR_ARG, "block %d: dims, scheme, RK order and viscosity must match block 0", i);
        if (c->dev != c0->dev) return set_err(SHX_ERR_ARG, "block %d is on another device", i);
    }
    const int dims = c0->gm.dims;
    shx_blocks *g = new shx_blocks();
    g->b.assign(blocks, blocks + nblocks);
    for (int i = 0; i < nblocks; i++)
        for (int f = 0; f < 6; f++) {
            const int kind = (f < 2 * dims) ? blocks[i]->d.bc[f] : SHX_BC_PERIODIC;
            g->kind[i][f] = (kind == SHX_BC_PERIODIC) ? XK_PERIODIC : (kind == SHX_BC_INTERFACE ? XK_INTERFACE : XK_WRITER);
            g->donor[i][f] = -1; g->dface[i][f] = -1;
            for (int q = 0; q < 3; q++) g->tr[i][f][q] = q + 1;
        }
    auto fail = [&](int code, const char *msg, int k) {
        delete g;
        return set_err(code, "connection %d: %s", k, msg);
    };
    for (int k = 0; k < nifc; k++) {
        const shx_interface &e = ifc[k];
        if (e.block < 0 || e.block >= nblocks || e.donor < 0 || e.donor >= nblocks) return fail(SHX_ERR_ARG, "block index out of range", k);
        if (e.face < 0 || e.face >= 2 * dims || e.donor_face < 0 || e.donor_face >= 2 * dims) return fail(SHX_ERR_ARG, "face out of range", k);
        if (g->kind[e.block][e.face] != XK_INTERFACE || g->kind[e.donor][e.donor_face] != XK_INTERFACE)
            return fail(SHX_ERR_ARG, "both faces must be of kind SHX_BC_INTERFACE", k);
        if (g->donor[e.block][e.face] >= 0) return fail(SHX_ERR_ARG, "the face is listed twice", k);
        // a signed permutation whose normal entry is the donor face's axis; whole faces
        int seen = 0;
        for (int q = 0; q < dims; q++) {
            const int t = e.transform[q], at = (t > 0 ? t : -t) - 1;
            if (at < 0 || at >= dims || (seen & (1 << at))) return fail(SHX_ERR_ARG, "transform is not a signed permutation", k);
            seen |= 1 << at;
        }
        const int a = e.face / 2, ad = e.donor_face / 2;
        if ((e.transform[a] > 0 ? e.transform[a] : -e.transform[a]) - 1 != ad)
            return fail(SHX_ERR_ARG, "the transform must take the face's normal to the donor face's", k);
        const shx_case *B = blocks[e.block], *D = blocks[e.donor];
        for (int q = 0; q < dims; q++) {
            if (q == a) continue;
            const int aq = (e.transform[q] > 0 ? e.transform[q] : -e.transform[q]) - 1;
            if (B->gm.n[q] != D->gm.n[aq]) return fail(SHX_ERR_UNSUPPORTED, "faces join only whole (tangential extents differ)", k);
        }
        if (B->gm.n[a] < B->gm.g || D->gm.n[ad] < B->gm.g) return fail(SHX_ERR_ARG, "block thinner than its ghost layers", k);
        g->donor[e.block][e.face] = e.donor;
        g->dface[e.block][e.face] = e.donor_face;
        for (int q = 0; q < 3; q++) g->tr[e.block][e.face][q] = (q < dims) ? e.transform[q] : q + 1;
    }
    for (int i = 0; i < nblocks; i++)
        for (int f = 0; f < 2 * dims; f++) {
            if (g->kind[i][f] != XK_INTERFACE) continue;
            const int d = g->donor[i][f], fd = g->dface[i][f];
            if (d < 0) { delete g; return set_err(SHX_ERR_ARG, "block %d face %d: interface without a connection", i, f); }
            if (g->donor[d][fd] != i || g->dface[d][fd] != f) {
                delete g;
                return set_err(SHX_ERR_ARG, "block %d face %d: the donor face does not connect back", i, f);
            }
            // the two sides' transforms are inverse permutations of each other
            for (int q = 0; q < dims; q++) {
                const int t = g->tr[i][f][q], aq = (t > 0 ? t : -t) - 1;
                const int back = g->tr[d][fd][aq], bq = (back > 0 ? back : -back) - 1;
                if (bq != q || (q != f / 2 && (t > 0) != (back > 0))) {
                    delete g;
                    return set_err(SHX_ERR_ARG, "block %d face %d: transforms of the two sides do not invert", i, f);
                }
            }
        }
    hipError_t e = hipStreamCreate(&g->stream);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_x, hipEventDisableTiming);
    g->ev_b.assign(nblocks, nullptr);
    for (int i = 0; i < nblocks && e == hipSuccess; i++) e = hipEventCreateWithFlags(&g->ev_b[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        shx_blocks_destroy(g);
        return set_err(SHX_ERR_HIP, "group setup: %s", hipGetErrorString(e));
    }
    for (int i = 0; i < nblocks; i++) blocks[i]->blocks = g;
    *out = g;
    return SHX_OK;
}

extern "C" int shx_blocks_destroy(shx_blocks *g)
{
    if (!g) return SHX_OK;
    if (g->stream) hipStreamSynchronize(g->stream);
    for (shx_case *c : g->b) {
        if (c->blocks == g) c->blocks = nullptr;
    }
    for (hipEvent_t ev : g->ev_b)
        if (ev) hipEventDestroy(ev);
    if (g->ev_x) hipEventDestroy(g->ev_x);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
    return SHX_OK;
}

// the group's stream after every block's stream, and every block's after the group's
static int join_blocks(shx_blocks *g)
{
    for (size_t i = 0; i < g->b.size(); i++) {
        HIPCHK(hipEventRecord(g->ev_b[i], g->b[i]->stream));
        HIPCHK(hipStreamWaitEvent(g->stream, g->ev_b[i], 0));
    }
    return SHX_OK;
}
static int fork_blocks(shx_blocks *g)
{
    HIPCHK(hipEventRecord(g->ev_x, g->stream));
    for (shx_case *c : g->b) HIPCHK(hipStreamWaitEvent(c->stream, g->ev_x, 0));
    return SHX_OK;
}

extern "C" int shx_blocks_transfer_metric(shx_blocks *g)
{
    if (!g) return set_err(SHX_ERR_ARG, "null group");
    for (shx_case *c : g->b)
        if (!c->has_metric) return set_err(SHX_ERR_STATE, "a block has no metric yet");
    int rc;
    if ((rc = join_blocks(g)) || (rc = gather(g, true)) || (rc = fork_blocks(g))) return rc;
    HIPCHK(hipStreamSynchronize(g->stream));
    return SHX_OK;
}

extern "C" int shx_blocks_run(shx_blocks *g, double tau, int nsteps)
{
    if (!g) return set_err(SHX_ERR_ARG, "null group");
    if (nsteps < 0) return set_err(SHX_ERR_ARG, "negative step count");
    int rc;
    for (shx_case *c : g->b)
        if ((rc = check_ready(c))) return rc;
    const int order = g->b[0]->d.time_order;
    for (int step = 0; step < nsteps; step++) {
        for (shx_case *c : g->b)
            if ((rc = step_begin(c))) return rc;
        for (int s = 0; s < order; s++) {
            if ((rc = join_blocks(g)) || (rc = gather(g, false)) || (rc = fork_blocks(g))) return rc;
            for (shx_case *c : g->b)
                if ((rc = step_stage(c, s, tau, STEP_NAN))) return rc;
        }
        for (shx_case *c : g->b) step_end(c);
        int bad = 0;
        for (shx_case *c : g->b) {
            int fresh = 0;
            if ((rc = nan_found(c, fresh))) return rc;
            bad += fresh;
        }
        if (bad > 0)
            return set_err(SHX_ERR_NAN, "checkNAN: %d interior nodes with a non-finite density after step %d of %d "
                           "(set to 1e6)", bad, step + 1, nsteps);
    }
    if ((rc = join_blocks(g))) return rc;
    return SHX_OK;
}

// ---- batches of independent cases --------------------------------------------------------------
// A batch advances many cases of uneven size together.  Cases that share the kernel templates
// (dimension, scheme, metric kind, arithmetic form, viscosity, RK order) form a group; per stage
// every kernel class of a group is one launch over the concatenated tile lists of its cases
// (k_*_batch), and the launches of one full RK step of every group are captured once into a HIP
// graph that is replayed per step.  The step keeps each case's U_lastStep buffer in place: the
// last stage writes the new state over U_n pointwise (each node reads U_n before it is
// overwritten), so the captured pointers stay valid from step to step.  Blocks of a launch walk
// its work list with a grid stride; the number of resident blocks per class is chosen by
// shx_batch_tune.
struct BatchLaunch {
    const void *fn, *fnp;          // flat and persistent kernels; launched: fn when grid == total
    const void *run;
    int threads, total, grid, klass;
    int group;                     // kernel-template group (one lane of the step graph)
    size_t args_off, work_off;
    const void *A;
    const int4 *work;
};

enum { BK_GHOST, BK_VALUES, BK_GRAD, BK_WENO, BK_VISC, BK_RK, BK_NUM };
static const char *g_bknames[BK_NUM] = {"ghost", "values", "grad", "weno", "visc", "rk"};

struct shx_batch {
    std::vector<shx_case *> cases;
    std::vector<double> tau;
    hipStream_t stream = nullptr;
    std::vector<BatchLaunch> seq;
    std::vector<char> host;      // argument tables, laid out as on the device
    char *dev = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    int factor[BK_NUM];          // resident blocks per CU per class (0: one block per work item)
    int ncu = 0;
    int use_graph = 1;           // 0: launch the step's kernels directly (SHX_BATCH_GRAPH=0)
    // Groups share no data, so every group runs on a lane stream of its own between a fork and a
    // join on `stream`: kernels of different groups (a memory-bound fill of one beside the sweeps
    // of another, the tail of one launch beside the head of the next) overlap on the device.
    int ngroups = 0, cur_group = 0;
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    // checkNAN fused into every case's last stage: running counts per case on the device, read by
    // shx_batch_synchronize into a pinned host copy; nan_seen: the counts already reported
    int *d_nan = nullptr, *h_nan = nullptr;
    std::vector<int> nan_seen;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// stage buffers with the state kept in place: (in, out) per stage
static void batch_buffers(const shx_case *c, int s, int &in, int &out)
{
    const int cur = c->cur, bx = (cur + 1) % 3, by = (cur + 2) % 3;
    if (c->d.time_order == 4) {
        static const int io[4][2] = {{0, 1}, {1, 2}, {2, 1}, {1, 0}};
        const int map[3] = {cur, bx, by};
        in = map[io[s][0]]; out = map[io[s][1]];
    } else {
        static const int io[3][2] = {{0, 1}, {1, 2}, {2, 0}};
        const int map[3] = {cur, bx, by};
        in = map[io[s][0]]; out = map[io[s][1]];
    }
}

// append one launch over `cs` with per-case argument records `recs` (rsize bytes each) and grids
// Work-list order, as the single-case launches order their tiles (sweep_tile): ORDER_NODES keeps
// every case's blocks in x order; ORDER_SWEEP lists sweep tiles fastest and places them on the
// XCDs as xcd_order does.
enum { ORDER_NODES, ORDER_SWEEP };
static int host_xcd_order(const int p, const int total)
{
    const int per = total >> 3, rem = total & 7, x = p & 7, s = p >> 3;
    return x * per + std::min(x, rem) + s;
}
static void batch_add(shx_batch *b, const void *fn, const void *fnp, int threads, int klass,
                      const std::vector<char> &recs, size_t rsize, const std::vector<dim3> &grids,
                      int order = ORDER_NODES)
{
    BatchLaunch L;
    memset(&L, 0, sizeof L);
    L.fn = fn; L.fnp = fnp; L.threads = threads; L.klass = klass; L.group = b->cur_group;
    const int nc = (int)grids.size();
    std::vector<int4> tiles;
    for (int i = 0; i < nc; i++) {
        const dim3 g = grids[i];      // node-parallel kernels: (blocks, 1, 1)
        if (order == ORDER_SWEEP) {
            for (unsigned y = 0; y < g.y; y++)
                for (unsigned x = 0; x < g.x; x++)
                    for (unsigned z = 0; z < g.z; z++) tiles.push_back(make_int4(i, (int)x, (int)y, (int)z));
        } else {
            for (unsigned z = 0; z < g.z; z++)
                for (unsigned y = 0; y < g.y; y++)
                    for (unsigned x = 0; x < g.x; x++) tiles.push_back(make_int4(i, (int)x, (int)y, (int)z));
        }
    }
    std::vector<int4> work(tiles.size());
    const int nt = (int)tiles.size();
    for (int q = 0; q < nt; q++) work[q] = (order == ORDER_NODES) ? tiles[q] : tiles[host_xcd_order(q, nt)];
    L.total = (int)work.size();
    if (work.empty()) work.push_back(make_int4(0, 0, 0, 0));
    L.args_off = align256(b->host.size());
    b->host.resize(L.args_off + recs.size());
    memcpy(b->host.data() + L.args_off, recs.data(), recs.size());
    L.work_off = align256(b->host.size());
    b->host.resize(L.work_off + work.size() * sizeof(int4));
    memcpy(b->host.data() + L.work_off, work.data(), work.size() * sizeof(int4));
    (void)rsize;
    b->seq.push_back(L);
}

template <class T>
static void put(std::vector<char> &recs, const T &r)
{
    const size_t o = recs.size();
    recs.resize(o + sizeof(T));
    memcpy(recs.data() + o, &r, sizeof(T));
}

// the launches of one RK step of one group (cases with identical kernel templates)
static int batch_group(shx_batch *b, const std::vector<shx_case *> &g, const std::vector<double> &tau)
{
    shx_case *c0 = g[0];
    const int dims = c0->gm.dims, S = c0->d.space_order;
    const int nc = (int)g.size();
    for (int s = 0; s < c0->d.time_order; s++) {
        std::vector<RkArgs> rk(nc);
        std::vector<double *const *> in(nc), out(nc);
        for (int i = 0; i < nc; i++) {
            int ib, ob;
            batch_buffers(g[i], s, ib, ob);
            in[i] = g[i]->prim[ib];
            out[i] = g[i]->prim[ob];
            rk[i] = rk_args(g[i], g[i]->prim[ib], g[i]->prim[ob], s, tau[i]);
            if (s == c0->d.time_order - 1 && b->d_nan) {
                // checkNAN on the step's new state (as shx_run), counted per case
                const long k = std::find(b->cases.begin(), b->cases.end(), g[i]) - b->cases.begin();
                rk[i].nan_cnt = b->d_nan + k;
                rk[i].nanP = g[i]->P;
            }
        }
        if (s == 0) {   // start of the step: step-start ghost state of every Rudy face
            for (int f = 0; f < 2 * dims; f++) {
                std::vector<char> recs;
                std::vector<dim3> grids;
                long any = 0;
                for (int i = 0; i < nc; i++) {
                    BcArgs a;
                    long nodes;
                    face_args(g[i], in[i], f, tau[i], a, nodes);
                    if (g[i]->d.bc[f] != SHX_BC_OUTFLOW_RUDY) nodes = 0;
                    put(recs, a);
                    grids.push_back(dim3(nblk(nodes, 256)));
                    any += nodes;
                }
                if (any) batch_add(b, (const void *)&k_face_save_batch<false>, (const void *)&k_face_save_batch<true>, 256, BK_GHOST, recs, sizeof(BcArgs), grids);
            }
        }
        if (c0->d.inviscid) {   // periodic ghost fill of the stage input
            std::vector<char> recs;
            std::vector<dim3> grids;
            for (int i = 0; i < nc; i++) {
                GhostArgs a;
                memset(&a, 0, sizeof a);
                for (int m = 0; m < 5; m++) a.f.f[m] = in[i][m];
                a.nf = 5; a.gm = g[i]->gm;
                put(recs, a);
                grids.push_back(dim3(nblk(g[i]->gm.ntot, 256)));
            }
            batch_add(b, (const void *)&k_ghost_batch<false>, (const void *)&k_ghost_batch<true>, 256, BK_GHOST, recs, sizeof(GhostArgs), grids);
        }
        // face writers, one launch per face in the reference's order (after CalcValues when viscous);
        // a case contributes no blocks to the launch of a periodic face
        auto add_faces = [&]() {
            for (int f = 0; f < 2 * dims; f++) {
                std::vector<char> recs;
                std::vector<dim3> grids;
                long any = 0;
                for (int i = 0; i < nc; i++) {
                    BcArgs a;
                    long nodes;
                    face_args(g[i], in[i], f, tau[i], a, nodes);
                    put(recs, a);
                    grids.push_back(dim3(nblk(nodes, 256)));
                    any += nodes;
                }
                if (any) batch_add(b, (const void *)&k_face_batch<false>, (const void *)&k_face_batch<true>, 256, BK_GHOST, recs, sizeof(BcArgs), grids);
            }
        };
        if (c0->d.inviscid) {
            // stored p and c of the cases with subsonic outflow faces
            std::vector<char> recs;
            std::vector<dim3> grids;
            long any = 0;
            for (int i = 0; i < nc; i++) {
                ValuesArgs a;
                memset(&a, 0, sizeof a);
                for (int m = 0; m < 5; m++) a.pr.f[m] = in[i][m];
                a.P = g[i]->P; a.C = g[i]->C; a.gm = g[i]->gm; a.k = g[i]->k;
                put(recs, a);
                const long nodes = g[i]->pc ? g[i]->gm.ntot : 0;
                grids.push_back(dim3(nblk(nodes, 256)));
                any += nodes;
            }
            if (any) batch_add(b, (const void *)&k_values_batch<false>, (const void *)&k_values_batch<true>, 256, BK_VALUES, recs, sizeof(ValuesArgs), grids);
            add_faces();
        }
        if (!c0->d.inviscid) {
            // periodic ghost fill of the stage input with CalcValues
            std::vector<char> recs, grecs;
            std::vector<dim3> grids, ggrids;
            for (int i = 0; i < nc; i++) {
                shx_case *c = g[i];
                GhostValuesArgs a;
                memset(&a, 0, sizeof a);
                for (int m = 0; m < 5; m++) a.pr.f[m] = in[i][m];
                a.T = c->T; a.mu = c->mu; a.P = c->P; a.C = c->C; a.gm = c->gm; a.k = c->k;
                put(recs, a);
                grids.push_back(dim3(nblk(c->gm.ntot, 256)));
                GradArgs ga;
                memset(&ga, 0, sizeof ga);
                for (int m = 0; m < 5; m++) ga.pr.f[m] = in[i][m];
                ga.T = c->T;
                for (int m = 0; m < 12; m++) ga.gr.f[m] = c->grad[m];
                ga.gm = c->gm; ga.k = c->k; ga.ci = 32; ga.tgrad = c->diag ? 0 : 1;
                put(grecs, ga);
                if (dims == 3) ggrids.push_back(dim3(nblk(c->gm.ng[2], GK), nblk(c->gm.ng[1], GJ), nblk(c->gm.ng[0], 32)));
                else ggrids.push_back(dim3(nblk(c->gm.ntot, 256)));
            }
            batch_add(b, (const void *)&k_ghost_values_batch<false>, (const void *)&k_ghost_values_batch<true>, 256, BK_VALUES, recs, sizeof(GhostValuesArgs), grids);
            add_faces();
            if (dims == 3) {
                // the march with the viscous xi terms follows the WENO xi sweep (below)
                if (!fuse_xi(c0))
                    batch_add(b, march_kernel(S, !c0->diag, true, false), march_kernel(S, !c0->diag, true, true), GK * GJ,
                              BK_GRAD, grecs, sizeof(GradArgs), ggrids, ORDER_NODES);
            } else {
                batch_add(b, (const void *)&k_grad2d_batch<false>, (const void *)&k_grad2d_batch<true>, 256, BK_GRAD, grecs, sizeof(GradArgs), ggrids);
            }
        }
        // the same sweep sequence as stage_fluxes (fused eta and zeta directions where they apply)
        const bool fz = fuse_zeta(c0), fx = fuse_xi(c0), fe = fuse_eta(c0);
        const int nsep = fz ? dims - 1 : dims;
        for (int dir = 0; dir < nsep; dir++) {
            if (fe && dir == 1) continue;
            const Pick p = weno_pick(c0, dir);
            std::vector<char> recs;
            std::vector<dim3> grids;
            for (int i = 0; i < nc; i++) {
                const WenoArgs a = weno_args(g[i], in[i], dir, dir == 0);
                put(recs, a);
                grids.push_back(sweep_grid(a.nA, a.oA, a.alignA, a.nB, a.nD, p));
            }
            batch_add(b, p.batch, p.batchp, p.TA * p.TDI, BK_WENO, recs, sizeof(WenoArgs), grids, ORDER_SWEEP);
        }
        if (fx) {
            std::vector<char> recs;
            std::vector<dim3> grids;
            for (int i = 0; i < nc; i++) {
                put(recs, grad_args(g[i], in[i], true));
                grids.push_back(march_grid(g[i]->gm, true));
            }
            batch_add(b, (const void *)&k_march_xi_batch<false>, (const void *)&k_march_xi_batch<true>, GKX * GJX,
                      BK_GRAD, recs, sizeof(GradArgs), grids, ORDER_NODES);
        }
        if (fz) {
            const Pick p = zeta_pick(c0);
            std::vector<char> zrecs;
            std::vector<dim3> grids;
            // viscous xi (unless the march took it) and eta (or the fused eta kernel), then the
            // fused zeta kernel
            for (int dir = fx ? 1 : 0; dir < nsep; dir++) {
                if (fe && dir == 1) {
                    const Pick pe = eta_pick(c0);
                    std::vector<char> recs;
                    std::vector<dim3> eg;
                    for (int i = 0; i < nc; i++) {
                        const EtaArgs z = eta_args(g[i], in[i]);
                        put(recs, z);
                        eg.push_back(sweep_grid(z.w.nA, z.w.oA, z.w.alignA, z.w.nB, z.w.nD, pe));
                    }
                    batch_add(b, pe.batch, pe.batchp, pe.TA * pe.TDI, BK_VISC, recs, sizeof(EtaArgs), eg, ORDER_SWEEP);
                    continue;
                }
                const Pick pv = visc_pick(c0, dir);
                std::vector<char> recs;
                std::vector<dim3> vg;
                for (int i = 0; i < nc; i++) {
                    const ViscArgs a = visc_args(g[i], in[i], dir, nullptr);
                    put(recs, a);
                    vg.push_back(sweep_grid(a.nA, a.oA, a.alignA, a.nB, a.nD, pv));
                }
                batch_add(b, pv.batch, pv.batchp, pv.TA * pv.TDI, BK_VISC, recs, sizeof(ViscArgs), vg,
                          sweep_order(true, dir) ? ORDER_SWEEP : ORDER_NODES);
            }
            for (int i = 0; i < nc; i++) {
                const ZetaArgs z = zeta_args(g[i], in[i], &rk[i]);
                put(zrecs, z);
                grids.push_back(sweep_grid(z.w.nA, z.w.oA, z.w.alignA, z.w.nB, z.w.nD, p));
            }
            batch_add(b, p.batch, p.batchp, p.TA * p.TDI, BK_VISC, zrecs, sizeof(ZetaArgs), grids, ORDER_SWEEP);
        } else if (!c0->d.inviscid) {
            for (int dir = 0; dir < dims; dir++) {
                const Pick p = visc_pick(c0, dir);
                std::vector<char> recs;
                std::vector<dim3> grids;
                for (int i = 0; i < nc; i++) {
                    const ViscArgs a = visc_args(g[i], in[i], dir, dir == dims - 1 ? &rk[i] : nullptr);
                    put(recs, a);
                    grids.push_back(sweep_grid(a.nA, a.oA, a.alignA, a.nB, a.nD, p));
                }
                batch_add(b, p.batch, p.batchp, p.TA * p.TDI, BK_VISC, recs, sizeof(ViscArgs), grids,
                          sweep_order(true, dir) ? ORDER_SWEEP : ORDER_NODES);
            }
        } else {
            std::vector<char> recs;
            std::vector<dim3> grids;
            for (int i = 0; i < nc; i++) {
                shx_case *c = g[i];
                UpdArgs A;
                memset(&A, 0, sizeof A);
                for (int m = 0; m < 5; m++) A.q.f[m] = c->q[m];
                A.J = c->met[9]; A.gm = c->gm; A.rk = rk[i];
                put(recs, A);
                grids.push_back(dim3(nblk((long)c->gm.n[0] * c->gm.n[1] * c->gm.n[2], 256)));
            }
            batch_add(b, (const void *)&k_rk_batch<false>, (const void *)&k_rk_batch<true>, 256, BK_RK, recs, sizeof(UpdArgs), grids);
        }
        {   // ghost layers of non-periodic directions into the next stage buffer (ghost nodes only,
            // so the in-place last stage, whose update writes interior nodes of U_n, is unaffected)
            std::vector<char> recs;
            std::vector<dim3> grids;
            long any = 0;
            for (int i = 0; i < nc; i++) {
                CarryArgs a;
                memset(&a, 0, sizeof a);
                for (int m = 0; m < 5; m++) { a.from.f[m] = in[i][m]; a.to.f[m] = out[i][m]; }
                a.gm = g[i]->gm;
                put(recs, a);
                const long nodes = g[i]->faces ? g[i]->gm.ntot : 0;
                grids.push_back(dim3(nblk(nodes, 256)));
                any += nodes;
            }
            if (any) batch_add(b, (const void *)&k_carry_batch<false>, (const void *)&k_carry_batch<true>, 256, BK_GHOST, recs, sizeof(CarryArgs), grids);
        }
    }
    return SHX_OK;
}

static bool same_group(const shx_case *a, const shx_case *c)
{
    return a->gm.dims == c->gm.dims && a->d.space_order == c->d.space_order && a->diag == c->diag &&
           a->d.precision == c->d.precision && a->d.inviscid == c->d.inviscid && a->d.time_order == c->d.time_order &&
           a->wtile[0] == c->wtile[0] && a->wtile[1] == c->wtile[1] && a->wtile[2] == c->wtile[2];
}

// (re)build the argument tables of every launch for the current time steps
static int batch_build(shx_batch *b)
{
    b->seq.clear();
    b->host.clear();
    b->ngroups = 0;
    std::vector<int> done(b->cases.size(), 0);
    for (size_t i = 0; i < b->cases.size(); i++) {
        if (done[i]) continue;
        std::vector<shx_case *> g;
        std::vector<double> t;
        for (size_t j = i; j < b->cases.size(); j++)
            if (!done[j] && same_group(b->cases[i], b->cases[j])) {
                g.push_back(b->cases[j]);
                t.push_back(b->tau[j]);
                done[j] = 1;
            }
        b->cur_group = b->ngroups++;
        int rc = batch_group(b, g, t);
        if (rc) return rc;
    }
    return SHX_OK;
}

static int batch_grids(shx_batch *b)
{
    for (BatchLaunch &L : b->seq) {
        const int f = b->factor[L.klass];
        if (f <= 0) { L.grid = std::max(L.total, 1); L.run = L.fn; continue; }
        int per = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, L.fnp, L.threads, 0));
        L.grid = std::max(1, std::min(L.total, std::max(1, per) * b->ncu * f));
        L.run = L.fnp;
    }
    return SHX_OK;
}

// lane streams and events for the current number of groups
static int batch_lanes(shx_batch *b)
{
    if (!b->fork) HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    while ((int)b->lanes.size() < b->ngroups) {
        hipStream_t s;
        hipEvent_t e;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        b->lanes.push_back(s);
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->joins.push_back(e);
    }
    return SHX_OK;
}

// one RK step of every group: fork from `stream`, each group's launches on its lane, join
static hipError_t batch_enqueue(shx_batch *b)
{
    const bool fan = b->ngroups > 1;
    hipError_t e = hipSuccess;
    if (fan) {
        e = hipEventRecord(b->fork, b->stream);
        for (int g = 0; g < b->ngroups && e == hipSuccess; g++) e = hipStreamWaitEvent(b->lanes[g], b->fork, 0);
    }
    for (BatchLaunch &L : b->seq) {
        if (e != hipSuccess) break;
        void *args[] = {&L.A, &L.work, &L.total};
        e = hipLaunchKernel(L.run, dim3(L.grid), dim3(L.threads), args, 0, fan ? b->lanes[L.group] : b->stream);
    }
    if (fan)
        for (int g = 0; g < b->ngroups && e == hipSuccess; g++) {
            e = hipEventRecord(b->joins[g], b->lanes[g]);
            if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->joins[g], 0);
        }
    return e;
}

static int batch_capture(shx_batch *b)
{
    if (b->exec) { hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    int rc = batch_lanes(b);
    if (rc) return rc;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = batch_enqueue(b);
    if (e != hipSuccess) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(b->stream, &g);
        if (g) hipGraphDestroy(g);
        return set_err(SHX_ERR_HIP, "batch capture: %s", hipGetErrorString(e));
    }
    HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return SHX_OK;
}

// upload the tables (same layout on every rebuild, so captured pointers stay valid)
static int batch_upload(shx_batch *b, bool fresh)
{
    if (fresh) {
        if (b->dev) hipFree(b->dev);
        b->dev = nullptr;
        HIPCHK(hipMalloc((void **)&b->dev, std::max<size_t>(b->host.size(), 256)));
    }
    HIPCHK(hipMemcpyAsync(b->dev, b->host.data(), b->host.size(), hipMemcpyHostToDevice, b->stream));
    for (BatchLaunch &L : b->seq) {
        L.A = b->dev + L.args_off;
        L.work = (const int4 *)(b->dev + L.work_off);
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return SHX_OK;
}

extern "C" int shx_batch_create(shx_case *const *cases, int ncases, const double *tau, shx_batch **out)
{
    if (!cases || !tau || !out || ncases <= 0) return set_err(SHX_ERR_ARG, "null argument or empty batch");
    *out = nullptr;
    for (int i = 0; i < ncases; i++) {
        shx_case *c = cases[i];
        int rc = check_ready(c);
        if (rc) return rc;
        if (!(tau[i] > 0.0)) return set_err(SHX_ERR_ARG, "case %d: tau must be positive", i);
        for (int j = 0; j < i; j++)
            if (cases[j] == c) return set_err(SHX_ERR_ARG, "case %d appears twice", i);
        if (c->batch) return set_err(SHX_ERR_STATE, "case %d is in another batch", i);
        if (c->iface) return set_err(SHX_ERR_UNSUPPORTED, "case %d has interface faces (a block of a multi-block case)", i);
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    shx_batch *b = new shx_batch();
    b->cases.assign(cases, cases + ncases);
    b->tau.assign(tau, tau + ncases);
    for (int k = 0; k < BK_NUM; k++) b->factor[k] = 0;
    {
        const char *env = getenv("SHX_BATCH_GRAPH");
        if (env && env[0] == '0') b->use_graph = 0;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipStreamCreate(&b->stream);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_nan, sizeof(int) * (size_t)ncases);
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_nan, sizeof(int) * (size_t)ncases, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_nan, 0, sizeof(int) * (size_t)ncases, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) { shx_batch_destroy(b); return set_err(SHX_ERR_HIP, "batch setup: %s", hipGetErrorString(e)); }
    b->nan_seen.assign((size_t)ncases, 0);
    int rc = batch_build(b);
    if (!rc) rc = batch_upload(b, true);
    if (!rc) rc = batch_grids(b);
    if (!rc) rc = batch_capture(b);
    if (rc) { shx_batch_destroy(b); return rc; }
    for (shx_case *c : b->cases) c->batch = b;
    *out = b;
    return SHX_OK;
}

extern "C" int shx_batch_set_tau(shx_batch *b, const double *tau)
{
    if (!b || !tau) return set_err(SHX_ERR_ARG, "null argument");
    for (size_t i = 0; i < b->cases.size(); i++)
        if (!(tau[i] > 0.0)) return set_err(SHX_ERR_ARG, "case %zu: tau must be positive", i);
    HIPCHK(hipStreamSynchronize(b->stream));
    b->tau.assign(tau, tau + b->cases.size());
    std::vector<BatchLaunch> old = b->seq;
    int rc = batch_build(b);
    if (rc) return rc;
    for (size_t i = 0; i < b->seq.size(); i++) { b->seq[i].grid = old[i].grid; b->seq[i].run = old[i].run; }
    return batch_upload(b, false);
}

extern "C" int shx_batch_run(shx_batch *b, int nsteps)
{
    if (!b) return set_err(SHX_ERR_ARG, "null batch");
    if (nsteps < 0) return set_err(SHX_ERR_ARG, "negative step count");
    // order the batch after earlier work on every case's stream, and later work after the batch
    for (shx_case *c : b->cases) {
        HIPCHK(hipEventRecord(b->done, c->stream));
        HIPCHK(hipStreamWaitEvent(b->stream, b->done, 0));
    }
    for (int s = 0; s < nsteps; s++) {
        if (b->use_graph) {
            HIPCHK(hipGraphLaunch(b->exec, b->stream));
        } else {
            HIPCHK(batch_enqueue(b));
        }
    }
    HIPCHK(hipEventRecord(b->done, b->stream));
    for (shx_case *c : b->cases) HIPCHK(hipStreamWaitEvent(c->stream, b->done, 0));
    return SHX_OK;
}

// Waits for the batch, then reads the fused checkNAN counts: SHX_ERR_NAN names the cases whose
// new states held a non-finite density in the steps since the last call (those nodes hold 1e6).
static int batch_nan_read(shx_batch *b)
{
    const size_t n = b->cases.size();
    HIPCHK(hipMemcpyAsync(b->h_nan, b->d_nan, sizeof(int) * n, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return SHX_OK;
}

extern "C" int shx_batch_synchronize(shx_batch *b)
{
    if (!b) return set_err(SHX_ERR_ARG, "null batch");
    int rc = batch_nan_read(b);
    if (rc) return rc;
    char msg[400];
    int len = 0, bad = 0;
    for (size_t i = 0; i < b->cases.size(); i++) {
        const int fresh = b->h_nan[i] - b->nan_seen[i];
        b->nan_seen[i] = b->h_nan[i];
        if (fresh <= 0) continue;
        if (bad++ < 8 && len < (int)sizeof msg - 64)
            len += snprintf(msg + len, sizeof msg - len, "%scase %zu: %d", bad > 1 ? ", " : "", i, fresh);
    }
    if (bad)
        return set_err(SHX_ERR_NAN, "checkNAN: interior nodes with a non-finite density (set to 1e6) in %d case%s "
                       "since the last synchronize (%s%s)", bad, bad > 1 ? "s" : "", msg, bad > 8 ? ", ..." : "");
    return SHX_OK;
}

extern "C" int shx_batch_nan_counts(shx_batch *b, int *counts)
{
    if (!b || !counts) return set_err(SHX_ERR_ARG, "null argument");
    int rc = batch_nan_read(b);
    if (rc) return rc;
    for (size_t i = 0; i < b->cases.size(); i++) counts[i] = b->h_nan[i];
    return SHX_OK;
}

// Autotuning: for each kernel class in turn, time whole replayed steps with each candidate number
// of resident blocks per CU (0 = one block per work item) and keep the fastest.  The case states
// are restored afterwards, so tuning leaves the solution unchanged.
extern "C" int shx_batch_tune(shx_batch *b, int steps, int factors_out[6])
{
    if (!b) return set_err(SHX_ERR_ARG, "null batch");
    if (steps <= 0) steps = 1;
    HIPCHK(hipStreamSynchronize(b->stream));
    // snapshot U_lastStep of every case
    std::vector<double *> snap;
    for (shx_case *c : b->cases) {
        HIPCHK(hipStreamSynchronize(c->stream));
        const size_t bytes = sizeof(double) * (size_t)c->gm.nalloc;
        for (int m = 0; m < 5; m++) {
            double *p = nullptr;
            HIPCHK(hipMalloc((void **)&p, bytes));
            HIPCHK(hipMemcpyAsync(p, c->prim[c->cur][m], bytes, hipMemcpyDeviceToDevice, b->stream));
            snap.push_back(p);
        }
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    static const int cand[4] = {1, 2, 4, 0};
    int rc = SHX_OK;
    for (int k = 0; k < BK_NUM && !rc; k++) {
        bool used = false;
        for (const BatchLaunch &L : b->seq) used |= (L.klass == k);
        if (!used) continue;
        float best = 1e30f;
        int bestf = b->factor[k];
        for (int ci = 0; ci < 4 && !rc; ci++) {
            b->factor[k] = cand[ci];
            if ((rc = batch_grids(b)) || (rc = batch_capture(b))) break;
            HIPCHK(hipGraphLaunch(b->exec, b->stream));      // warm-up
            HIPCHK(hipEventRecord(e0, b->stream));
            for (int s = 0; s < steps; s++) HIPCHK(hipGraphLaunch(b->exec, b->stream));
            HIPCHK(hipEventRecord(e1, b->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; bestf = cand[ci]; }
        }
        b->factor[k] = bestf;
    }
    // restore the states
    size_t si = 0;
    for (shx_case *c : b->cases) {
        const size_t bytes = sizeof(double) * (size_t)c->gm.nalloc;
        for (int m = 0; m < 5; m++, si++) {
            HIPCHK(hipMemcpyAsync(c->prim[c->cur][m], snap[si], bytes, hipMemcpyDeviceToDevice, b->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    for (double *p : snap) hipFree(p);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    // the tuning steps are not part of the solution: their checkNAN counts are not reported
    if ((rc = batch_nan_read(b))) return rc;
    for (size_t i = 0; i < b->cases.size(); i++) b->nan_seen[i] = b->h_nan[i];
    if ((rc = batch_grids(b)) || (rc = batch_capture(b))) return rc;
    if (factors_out)
        for (int k = 0; k < BK_NUM; k++) factors_out[k] = b->factor[k];
    return SHX_OK;
}

extern "C" int shx_batch_destroy(shx_batch *b)
{
    if (!b) return SHX_OK;
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->dev) hipFree(b->dev);
    if (b->done) hipEventDestroy(b->done);
    for (hipStream_t l : b->lanes) { hipStreamSynchronize(l); hipStreamDestroy(l); }
    for (hipEvent_t e : b->joins) hipEventDestroy(e);
    if (b->fork) hipEventDestroy(b->fork);
    if (b->d_nan) hipFree(b->d_nan);
    if (b->h_nan) hipHostFree(b->h_nan);
    if (b->stream) hipStreamDestroy(b->stream);
    for (shx_case *c : b->cases)
        if (c->batch == b) c->batch = nullptr;
    delete b;
    return SHX_OK;
}

// Tile autotuning: for every direction, time `reps` WENO sweeps of the current state with each
// tile variant and keep the fastest.  The sweeps only write Q, which every stage recomputes, so
// the state is unchanged.
extern "C" int shx_tune(shx_case *c, int reps, int choice[3])
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (reps <= 0) reps = 3;
    double *const *pr = c->prim[c->cur];
    if ((rc = ghost_fill(c, pr, 5))) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    for (int dir = 0; dir < c->gm.dims && !rc; dir++) {
        // the zeta sweep of a fusable case stays on the 4 x 64 tile of the fused kernel (k_zeta)
        if (dir == 2 && c->gm.dims == 3 && c->diag && !c->d.inviscid) { c->wtile[2] = 0; continue; }
        if (dir == 1 && fuse_eta(c)) { c->wtile[1] = 1; continue; }   // the fused eta kernel's 8 x 32 tile
        float best = 1e30f;
        int bestv = 0;
        for (int v = 0; v < NWTILE && !rc; v++) {
            c->wtile[dir] = v;
            if ((rc = weno_sweep(c, pr, dir, 1))) break;          // warm-up
            HIPCHK(hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps && !rc; r++) rc = weno_sweep(c, pr, dir, 1);
            HIPCHK(hipEventRecord(e1, c->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; bestv = v; }
        }
        c->wtile[dir] = bestv;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    if (choice)
        for (int d = 0; d < 3; d++) choice[d] = c->wtile[d];
    if (c->timing) drain_events(c);
    return SHX_OK;
}

extern "C" int shx_run(shx_case *c, double tau, int nsteps)
{
    if (nsteps < 0) return set_err(SHX_ERR_ARG, "negative step count");
    for (int s = 0; s < nsteps; s++) {
        int rc = step_rk(c, tau, STEP_NAN);
        if (rc) return rc;
        int bad = 0;
        if ((rc = nan_found(c, bad))) return rc;
        if (bad > 0)
            return set_err(SHX_ERR_NAN, "checkNAN: %d interior nodes with a non-finite density after step %d of %d "
                           "(set to 1e6)", bad, s + 1, nsteps);
        if (c->timing && c->ev_used > 4096) { rc = drain_events(c); if (rc) return rc; }
    }
    return SHX_OK;
}

// the maximum node CFL sum the last reduction into c->d_red left (its slots copied to h_io->red,
// stream synchronized)
static double red_max(shx_case *c)
{
    unsigned long long mb = 0;
    for (int s = 0; s < RED_SLOTS; s++) mb = std::max(mb, c->h_io->red[s]);
    double mx;
    memcpy(&mx, &mb, sizeof mx);
    return mx;
}
static int red_clear(shx_case *c)
{
    HIPCHK(hipMemsetAsync(c->d_red, 0, sizeof(unsigned long long) * RED_SLOTS, c->stream));
    return SHX_OK;
}
static int red_fetch(shx_case *c)
{
    HIPCHK(hipMemcpyAsync(c->h_io->red, c->d_red, sizeof(unsigned long long) * RED_SLOTS, hipMemcpyDeviceToHost,
                          c->stream));
    return SHX_OK;
}

extern "C" int shx_run_cfl(shx_case *c, double cfl, int nsteps, double *time)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (nsteps < 0 || !(cfl > 0.0)) return set_err(SHX_ERR_ARG, "negative step count or bad cfl");
    if (c->rudy) return set_err(SHX_ERR_UNSUPPORTED, "CFL stepping with Rudy outflow faces (their relaxation uses tau)");
    const Geom &gm = c->gm;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    // the first step's maximum from the current state; later ones from the fused reduction
    CFields5 pf;
    CFields10 mf;
    for (int m = 0; m < 5; m++) pf.f[m] = c->prim[c->cur][m];
    for (int m = 0; m < 10; m++) mf.f[m] = c->met[m];
    if ((rc = red_clear(c))) return rc;
    if ((rc = tbeg(c, KT_REDUCE))) return rc;
    k_cfl_max<<<nblk(nint, 256), 256, 0, c->stream>>>(pf, mf, gm, c->k, c->d_red);
    LAUNCH_CHECK();
    if ((rc = tend(c))) return rc;
    if ((rc = red_fetch(c))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    double t = 0.0;
    for (int s = 0; s < nsteps; s++) {
        const double mx = red_max(c);
        if (!(mx > 0.0) || !std::isfinite(mx)) return set_err(SHX_ERR_NAN, "CFL maximum %g before step %d", mx, s + 1);
        const double dt = cfl / mx;
        if ((rc = red_clear(c))) return rc;
        if ((rc = step_rk(c, dt, STEP_NAN | STEP_CFL))) return rc;
        if ((rc = red_fetch(c))) return rc;
        int bad = 0;
        if ((rc = nan_found(c, bad))) return rc;
        t += dt;
        if (time) *time = t;
        if (bad > 0)
            return set_err(SHX_ERR_NAN, "checkNAN: %d interior nodes with a non-finite density after step %d of %d "
                           "(set to 1e6)", bad, s + 1, nsteps);
        if (c->timing && c->ev_used > 4096) { rc = drain_events(c); if (rc) return rc; }
    }
    if (time) *time = t;
    return SHX_OK;
}

// The stored pressure of U_lastStep (the last stage's CalcValues value, which the RK update does
// not renew) is what the reference's post-processing reads.  Cases with subsonic outflow or slip
// wall faces keep it anyway; others start keeping it here, from the current state's own p.
static int keep_pressure(shx_case *c)
{
    if (c->pc) return SHX_OK;
    const size_t bytes = sizeof(double) * (size_t)c->gm.nalloc;
    HIPCHK(hipMalloc((void **)&c->P, bytes));
    HIPCHK(hipMalloc((void **)&c->C, bytes));
    CFields5 cp;
    for (int m = 0; m < 5; m++) cp.f[m] = c->prim[c->cur][m];
    k_calc_values<<<nblk(c->gm.ntot, 256), 256, 0, c->stream>>>(cp, nullptr, nullptr, c->P, c->C, c->gm, c->k);
    LAUNCH_CHECK();
    c->pc = 1;
    return SHX_OK;
}

// U_lastStep's ghosts in the reference are those of the last stage's input (filled at that stage's
// start); the state buffer gets them before it is read whole (post-processing, backups).  The next
// step refills the periodic ghosts and carries the others, so stepping is unaffected.
static int sync_ghosts(shx_case *c)
{
    if (c->last_in < 0 || c->last_in == c->cur) return SHX_OK;
    CFields5 a;
    Fields5 b;
    for (int m = 0; m < 5; m++) { a.f[m] = c->prim[c->last_in][m]; b.f[m] = c->prim[c->cur][m]; }
    k_copy_ghosts<<<nblk(c->gm.ntot, 256), 256, 0, c->stream>>>(a, b, c->gm);
    LAUNCH_CHECK();
    c->last_in = -1;
    return SHX_OK;
}

static int post_compute(shx_case *c)
{
    int rc = keep_pressure(c);
    if (rc) return rc;
    if ((rc = sync_ghosts(c))) return rc;
    const Geom &gm = c->gm;
    const size_t bytes = sizeof(double) * (size_t)gm.nalloc;
    for (int m = 0; m < 3; m++)
        if (!c->post[m]) {
            HIPCHK(hipMalloc((void **)&c->post[m], bytes));
            HIPCHK(hipMemsetAsync(c->post[m], 0, bytes, c->stream));
        }
    CFields5 cp;
    CFields10 mf;
    Fields5 out;
    memset(&out, 0, sizeof out);
    for (int m = 0; m < 5; m++) cp.f[m] = c->prim[c->cur][m];
    for (int m = 0; m < 10; m++) mf.f[m] = c->met[m];
    for (int m = 0; m < 3; m++) out.f[m] = c->post[m];
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    k_post<<<nblk(nint, 256), 256, 0, c->stream>>>(cp, c->P, mf, out, gm, c->k, c->d.space_order);
    LAUNCH_CHECK();
    return SHX_OK;
}

extern "C" int shx_post(shx_case *c, double *const out[3])
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!out) return set_err(SHX_ERR_ARG, "null output");
    if ((rc = post_compute(c))) return rc;
    for (int m = 0; m < 3; m++)
        if (out[m]) { rc = copy_out(c, out[m], c->post[m]); if (rc) return rc; }
    HIPCHK(hipStreamSynchronize(c->stream));
    return SHX_OK;
}

extern "C" int shx_film_open(shx_case *c, int samples)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (samples < 1) return set_err(SHX_ERR_ARG, "samples must be positive");
    if ((rc = keep_pressure(c))) return rc;
    const long nint = (long)c->gm.n[0] * c->gm.n[1] * c->gm.n[2];
    hipFree(c->film);
    c->film = nullptr;
    HIPCHK(hipMalloc((void **)&c->film, sizeof(double) * (size_t)8 * nint * samples));
    c->film_cap = samples;
    c->film_n = 0;
    c->film_time.assign(samples, 0.0);
    HIPCHK(hipStreamSynchronize(c->stream));
    return SHX_OK;
}

// one sample of the current state into the film (CalcValuesForPost + WriteValuesFromUToFilm)
static int film_sample(shx_case *c, double time)
{
    if (c->film_n >= c->film_cap) return set_err(SHX_ERR_STATE, "film full (%d samples)", c->film_cap);
    int rc = post_compute(c);
    if (rc) return rc;
    const Geom &gm = c->gm;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    CFields5 cp, po;
    memset(&po, 0, sizeof po);
    for (int m = 0; m < 5; m++) cp.f[m] = c->prim[c->cur][m];
    for (int m = 0; m < 3; m++) po.f[m] = c->post[m];
    k_film<<<nblk(nint, 256), 256, 0, c->stream>>>(cp, c->P, po, c->film, nint, c->film_n, gm);
    LAUNCH_CHECK();
    c->film_time[c->film_n++] = time;
    return SHX_OK;
}

extern "C" int shx_film_sample(shx_case *c, double time)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!c->film) return set_err(SHX_ERR_STATE, "no film: call shx_film_open");
    return film_sample(c, time);
}

extern "C" int shx_film_download(shx_case *c, int sample, double *const out[8], double *time)
{
    if (!c || !out) return set_err(SHX_ERR_ARG, "null argument");
    if (!c->film || sample < 0 || sample >= c->film_n) return set_err(SHX_ERR_ARG, "no such sample");
    const long nint = (long)c->gm.n[0] * c->gm.n[1] * c->gm.n[2];
    for (int m = 0; m < 8; m++)
        if (out[m])
            HIPCHK(hipMemcpyAsync(out[m], c->film + ((long)sample * 8 + m) * nint, sizeof(double) * nint,
                                  hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (time) *time = c->film_time[sample];
    return SHX_OK;
}

extern "C" int shx_history_download(shx_case *c, double *time, double *p, double *u, double *v, double *w)
{
    if (!c) return set_err(SHX_ERR_ARG, "null argument");
    if (!c->hist) return set_err(SHX_ERR_STATE, "no probe history: run shx_simulate with probes");
    const size_t fs = sizeof(double) * (size_t)c->hist_np * c->hist_iters;
    double *out[4] = {p, u, v, w};
    for (int m = 0; m < 4; m++)
        if (out[m]) HIPCHK(hipMemcpyAsync(out[m], c->hist + (size_t)m * c->hist_np * c->hist_iters, fs,
                                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (time) memcpy(time, c->hist_time.data(), sizeof(double) * c->hist_iters);
    return SHX_OK;
}

extern "C" int shx_mms_error(shx_case *c, double out[4])
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!out) return set_err(SHX_ERR_ARG, "null output");
    if (!c->has_bs) return set_err(SHX_ERR_STATE, "no exact state: call shx_upload_boundary_state");
    if ((rc = keep_pressure(c))) return rc;
    const Geom &gm = c->gm;
    const long nplane = (long)gm.n[0] * gm.n[1];
    const int nb = (int)nblk(nplane, 256);
    double *part = nullptr, *dout = nullptr;
    HIPCHK(hipMalloc((void **)&part, sizeof(double) * 4 * (size_t)nb));
    HIPCHK(hipMalloc((void **)&dout, sizeof(double) * 4));
    CFields5 cp;
    for (int m = 0; m < 5; m++) cp.f[m] = c->prim[c->cur][m];
    const double N = (double)gm.n[0] * gm.n[1] * gm.n[2];
    k_mms_error<<<nb, 256, 0, c->stream>>>(cp, c->P, c->bs[0], c->bs[5], gm, N, part);
    LAUNCH_CHECK();
    k_mms_error_total<<<1, 256, 0, c->stream>>>(part, nb, dout);
    LAUNCH_CHECK();
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    hipFree(part);
    hipFree(dout);
    return SHX_OK;
}

// ---- the entry point that allocates nothing -------------------------------------------------
// shx_mms_error_static: shx_mms_error with its partial sums and result in a file-scope device
// array instead of per-call allocations, for callers that cannot allocate (code under HIP graph
// capture, serving loops).  Every call on a device shares that device's array: calls that overlap
// on one device race.  No element needs clearing: k_mms_error writes every partial it later sums
// and k_mms_error_total writes every result.
constexpr int SCRATCH_BLOCKS = 4096;                   // n0 * n1 <= SCRATCH_BLOCKS * 256 interior nodes
static __device__ __align__(256) double scratch[4 + 4 * SCRATCH_BLOCKS];
constexpr int MAX_DEVICES = 64;
static double *g_scratch[MAX_DEVICES];                 // scratch's address on each device

// scratch's device address on the case's device, looked up once per device with that device
// current (in host code the array's name is only a stand-in for the symbol)
static int scratch_ptr(const shx_case *c, double **ptr)
{
    if (c->dev < 0 || c->dev >= MAX_DEVICES) return set_err(SHX_ERR_UNSUPPORTED, "device %d", c->dev);
    if (!g_scratch[c->dev]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != c->dev) HIPCHK(hipSetDevice(c->dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != c->dev) HIPCHK(hipSetDevice(cur));
        if (e != hipSuccess || !p) return set_err(SHX_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch[c->dev] = (double *)p;
    }
    *ptr = g_scratch[c->dev];
    return SHX_OK;
}

extern "C" int shx_mms_error_static(shx_case *c, double out[4])
{
    if (!c || !out) return set_err(SHX_ERR_ARG, "null argument");
    const Geom &gm = c->gm;
    const long nplane = (long)gm.n[0] * gm.n[1];
    const long nb = (long)nblk(nplane, 256);
    if (nb > SCRATCH_BLOCKS)
        return set_err(SHX_ERR_ARG, "shx_mms_error_static: n0 * n1 = %ld interior nodes per k plane exceeds the "
                       "static scratch (at most %d)", nplane, SCRATCH_BLOCKS * 256);
    int rc = check_ready(c);
    if (rc) return rc;
    if (!c->has_bs) return set_err(SHX_ERR_STATE, "no exact state: call shx_upload_boundary_state");
    if (!c->pc)
        return set_err(SHX_ERR_STATE, "shx_mms_error_static reads the stored pressure, which this case does not "
                       "keep yet (keeping it allocates): call shx_mms_error, shx_post or shx_film_open once first");
    double *scr = nullptr;
    if ((rc = scratch_ptr(c, &scr))) return rc;
    double *const dout = scr, *const part = scr + 4;
    CFields5 cp;
    for (int m = 0; m < 5; m++) cp.f[m] = c->prim[c->cur][m];
    const double N = (double)gm.n[0] * gm.n[1] * gm.n[2];
    k_mms_error<<<(unsigned)nb, 256, 0, c->stream>>>(cp, c->P, c->bs[0], c->bs[5], gm, N, part);
    LAUNCH_CHECK();
    k_mms_error_total<<<1, 256, 0, c->stream>>>(part, (int)nb, dout);
    LAUNCH_CHECK();
    HIPCHK(hipMemcpyAsync(c->h_io->mms, dout, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int m = 0; m < 4; m++) out[m] = c->h_io->mms[m];
    return SHX_OK;
}

extern "C" int shx_film_count(const shx_case *c, int *count)
{
    if (!c || !count) return set_err(SHX_ERR_ARG, "null argument");
    *count = c->film_n;
    return SHX_OK;
}

// state (and stored p) of the case into / out of backup slot b (WriteValuesFromU1ToU2,
// src/Functions.c:2484-2508: every padded node); a reload writes every stage buffer, as an upload does
static int backup_io(shx_case *c, int b, bool save)
{
    const size_t bytes = sizeof(double) * (size_t)c->gm.nalloc;
    if (save) { int rc = sync_ghosts(c); if (rc) return rc; }
    if (!c->bk[b][0])
        for (int m = 0; m < 5; m++) HIPCHK(hipMalloc((void **)&c->bk[b][m], bytes));
    if (c->pc && !c->bkP[b]) HIPCHK(hipMalloc((void **)&c->bkP[b], bytes));
    for (int m = 0; m < 5; m++) {
        if (save) {
            HIPCHK(hipMemcpyAsync(c->bk[b][m], c->prim[c->cur][m], bytes, hipMemcpyDeviceToDevice, c->stream));
        } else {
            for (int s = 0; s < 3; s++)
                HIPCHK(hipMemcpyAsync(c->prim[s][m], c->bk[b][m], bytes, hipMemcpyDeviceToDevice, c->stream));
        }
    }
    if (c->pc) {
        if (save) HIPCHK(hipMemcpyAsync(c->bkP[b], c->P, bytes, hipMemcpyDeviceToDevice, c->stream));
        else HIPCHK(hipMemcpyAsync(c->P, c->bkP[b], bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    if (!save) c->last_in = -1;
    return SHX_OK;
}

extern "C" int shx_simulate(shx_case *c, const shx_sim *sim, shx_sim_report *rep)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!sim || !rep) return set_err(SHX_ERR_ARG, "null argument");
    if (!(sim->tau > 0.0) || !(sim->L0 > 0.0) || !(sim->u0 > 0.0) || sim->end_iteration < sim->start_iteration ||
        sim->distance_backward < 1 || sim->number_resets < 0 || !(sim->decel_factor > 0.0) ||
        (sim->accelerator && !(sim->accel_factor > 0.0)))
        return set_err(SHX_ERR_ARG, "bad simulation parameters");
    const bool sampling = sim->samples > 0;
    if (sampling && (sim->iterations_between_samples < 1 || !c->film || c->film_cap < sim->samples))
        return set_err(SHX_ERR_ARG, "sampling needs iterations_between_samples >= 1 and a film of that many samples");
    if (sim->nprobes < 0 || (sim->nprobes > 0 && !sim->probes)) return set_err(SHX_ERR_ARG, "bad probe list");
    const Geom &gm = c->gm;
    // probes are the reference's linear indices (i ng1 ng2 + j ng2 + k); the device arrays use
    // padded row pitches
    std::vector<long> dev_probes(sim->nprobes);
    for (int q = 0; q < sim->nprobes; q++) {
        const long n = sim->probes[q];
        if (n < 0 || n >= gm.ntot) return set_err(SHX_ERR_ARG, "probe %d outside the grid", q);
        const int ijk[3] = {(int)(n / ((long)gm.ng[1] * gm.ng[2])), (int)((n / gm.ng[2]) % gm.ng[1]), (int)(n % gm.ng[2])};
        for (int d = 0; d < gm.dims; d++)
            if (ijk[d] < gm.lo[d] || ijk[d] >= gm.lo[d] + gm.n[d]) return set_err(SHX_ERR_ARG, "probe %d is not an interior node", q);
        dev_probes[q] = (long)ijk[0] * gm.st[0] + (long)ijk[1] * gm.st[1] + (long)ijk[2] * gm.st[2];
    }
    memset(rep, 0, sizeof *rep);
    if ((sampling || sim->nprobes > 0) && (rc = keep_pressure(c))) return rc;
    const int iters = sim->end_iteration - sim->start_iteration;
    if (sim->nprobes > 0) {
        hipFree(c->d_probes); hipFree(c->hist);
        c->d_probes = nullptr; c->hist = nullptr;
        HIPCHK(hipMalloc((void **)&c->d_probes, sizeof(long) * sim->nprobes));
        HIPCHK(hipMemcpy(c->d_probes, dev_probes.data(), sizeof(long) * sim->nprobes, hipMemcpyHostToDevice));
        if (iters > 0) {
            HIPCHK(hipMalloc((void **)&c->hist, sizeof(double) * 4 * (size_t)sim->nprobes * iters));
            HIPCHK(hipMemsetAsync(c->hist, 0, sizeof(double) * 4 * (size_t)sim->nprobes * iters, c->stream));
        }
        c->hist_np = sim->nprobes;
        c->hist_iters = iters;
        c->hist_time.assign(iters, 0.0);
    }
    const double tau_start = sim->tau, L0 = sim->L0, u0 = sim->u0;
    const int end = sim->end_iteration, dback = sim->distance_backward, dnan = sim->distance_nan;
    double tau = tau_start, time = sim->time, last = sim->time_last_action;
    int dfwd = sim->distance_forward;
    // slot 0 is the reference's backup1, slot 1 its backup2 (src/SHOCK.c:531-540)
    int it_b[2] = {1, 999999};
    double time_b[2] = {time, 0.0};
    // sampling counters (int_iterationCounterSamples, int_actualSample) and their backups
    int counter = 0, sample = c->film_n;
    int counter_b[2] = {0, 0}, sample_b[2] = {sample, 0};
    if ((rc = backup_io(c, 0, true))) return rc;
    int last_run = sim->start_iteration;
    for (int it = sim->start_iteration + 1; it <= end; it++) {
        last_run = it;
        time += tau * L0 / u0;
        if ((rc = step_rk(c, tau, STEP_NAN))) return rc;
        int bad = 0;
        if ((rc = nan_found(c, bad))) return rc;
        if (bad > 0) {
            rep->nan_steps++;
            if (rep->resets >= sim->number_resets) { rep->aborted = 1; break; }
            rep->resets++;
            tau *= sim->decel_factor;
            dfwd = sim->distance_forward;
            const int b = (it_b[0] < it_b[1]) ? 0 : 1;
            if ((rc = backup_io(c, b, false))) return rc;
            last = time;
            time = time_b[b];
            it = it_b[b];
            sample = sample_b[b];
            counter = counter_b[b];
        } else if ((time - last) > dnan * tau * L0 / u0 && sim->accelerator && it + dback < end && tau < tau_start) {
            dfwd = sim->distance_forward;
            tau = tau_start;
            last = time;
            rep->tau_resets++;
        } else if ((time - last) > dfwd * tau * L0 / u0 && sim->accelerator && it + dback < end && tau >= tau_start) {
            dfwd = (int)(dfwd * 4.);
            tau *= sim->accel_factor;
            last = time;
            rep->accelerations++;
        }
        if (it % dback == 0) {
            if ((rc = backup_io(c, 0, true))) return rc;
            time_b[0] = time; it_b[0] = it; sample_b[0] = sample; counter_b[0] = counter;
        } else if ((it + dback / 2) % dback == 0) {
            if ((rc = backup_io(c, 1, true))) return rc;
            time_b[1] = time; it_b[1] = it; sample_b[1] = sample; counter_b[1] = counter;
        }
        // sampling (src/SHOCK.c:694-717): every iterations_between_samples iterations, from
        // start_sampling on, while samples remain; a rollback rewinds the sample index
        counter++;
        if (sampling && counter % sim->iterations_between_samples == 0 && sample < sim->samples &&
            it >= sim->start_sampling) {
            c->film_n = sample;
            if ((rc = film_sample(c, time))) return rc;
            sample++;
        }
        // history row of this iteration (src/Functions.c:5381-5417 index the history by
        // iteration - (start + 1)); a rollback to a backup older than the start (the first backup
        // is iteration 1 whatever the start, src/SHOCK.c:531-540) gives iterations before the
        // first row, which the reference writes out of bounds: they are not recorded here
        const int row = it - (sim->start_iteration + 1);
        if (sim->nprobes > 0 && row >= 0 && row < iters) {
            CFields5 cp;
            for (int m = 0; m < 5; m++) cp.f[m] = c->prim[c->cur][m];
            k_probe<<<nblk(sim->nprobes, 64), 64, 0, c->stream>>>(cp, c->P, c->d_probes, sim->nprobes, c->hist, row, iters);
            LAUNCH_CHECK();
            c->hist_time[row] = time;
        }
        if (c->timing && c->ev_used > 4096) { rc = drain_events(c); if (rc) return rc; }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    rep->iteration = last_run;
    rep->time = time;
    rep->tau = tau;
    if (sampling) c->film_n = sample;
    rep->samples = sampling ? sample : c->film_n;
    return SHX_OK;
}

extern "C" int shx_rhs(shx_case *c, double *const q[5])
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!q) return set_err(SHX_ERR_ARG, "null output");
    if (c->iface) return set_err(SHX_ERR_UNSUPPORTED, "the case has interface faces (its ghosts need its multi-block case)");
    const Geom &gm = c->gm;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    // a Rudy face takes the current state as its step-start state, with tau = 0
    if ((rc = rudy_save(c, c->prim[c->cur]))) return rc;
    if ((rc = stage_fluxes(c, c->prim[c->cur], nullptr, 0.0))) return rc;
    if (c->has_src) {
        Fields5 qf;
        CFields5 sf;
        for (int m = 0; m < 5; m++) { qf.f[m] = c->q[m]; sf.f[m] = c->src[m]; }
        k_add_source<<<nblk(nint, 256), 256, 0, c->stream>>>(qf, sf, c->met[9], gm);
        LAUNCH_CHECK();
    }
    for (int m = 0; m < 5; m++) { rc = copy_out(c, q[m], c->q[m]); if (rc) return rc; }
    HIPCHK(hipStreamSynchronize(c->stream));
    return SHX_OK;
}

extern "C" int shx_check_metric(shx_case *c, int *negative)
{
    if (!c || !negative) return set_err(SHX_ERR_ARG, "null argument");
    if (!c->has_metric) return set_err(SHX_ERR_STATE, "no metric yet");
    HIPCHK(hipMemsetAsync(c->d_cnt, 0, sizeof(int), c->stream));
    k_check_metric<<<nblk(c->gm.ntot, 256), 256, 0, c->stream>>>(c->met[9], c->gm, c->d_cnt);
    LAUNCH_CHECK();
    HIPCHK(hipMemcpyAsync(negative, c->d_cnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SHX_OK;
}

extern "C" int shx_check_nan(shx_case *c, int *count)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!count) return set_err(SHX_ERR_ARG, "null output");
    const Geom &gm = c->gm;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    Fields5 pf;
    for (int m = 0; m < 5; m++) pf.f[m] = c->prim[c->cur][m];
    HIPCHK(hipMemsetAsync(c->d_cnt, 0, sizeof(int), c->stream));
    if ((rc = tbeg(c, KT_REDUCE))) return rc;
    k_check_nan<<<nblk(nint, 256), 256, 0, c->stream>>>(pf, c->P, gm, c->d_cnt);
    LAUNCH_CHECK();
    if ((rc = tend(c))) return rc;
    HIPCHK(hipMemcpyAsync(count, c->d_cnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SHX_OK;
}

extern "C" int shx_cfl_dt(shx_case *c, double cfl, double *dt)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!dt || !(cfl > 0.0)) return set_err(SHX_ERR_ARG, "bad cfl or null output");
    const Geom &gm = c->gm;
    const long nint = (long)gm.n[0] * gm.n[1] * gm.n[2];
    CFields5 pf;
    CFields10 mf;
    for (int m = 0; m < 5; m++) pf.f[m] = c->prim[c->cur][m];
    for (int m = 0; m < 10; m++) mf.f[m] = c->met[m];
    if ((rc = red_clear(c))) return rc;
    if ((rc = tbeg(c, KT_REDUCE))) return rc;
    k_cfl_max<<<nblk(nint, 256), 256, 0, c->stream>>>(pf, mf, gm, c->k, c->d_red);
    LAUNCH_CHECK();
    if ((rc = tend(c))) return rc;
    if ((rc = red_fetch(c))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    const double mx = red_max(c);
    *dt = mx > 0.0 ? cfl / mx : 0.0;
    return SHX_OK;
}

extern "C" int shx_copy_probe(shx_case *c, int reps)
{
    if (!c || reps < 0) return set_err(SHX_ERR_ARG, "bad argument");
    for (int r = 0; r < reps; r++) {
        k_copy<<<nblk(c->gm.nalloc, 256), 256, 0, c->stream>>>(c->T, c->mu, c->gm.nalloc);
        LAUNCH_CHECK();
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return SHX_OK;
}

extern "C" int shx_synchronize(shx_case *c)
{
    if (!c) return set_err(SHX_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(c->stream));
    return SHX_OK;
}

extern "C" int shx_set_timing(shx_case *c, int enable)
{
    if (!c) return set_err(SHX_ERR_ARG, "null handle");
    int rc = drain_events(c);
    if (rc) return rc;
    c->timing = enable ? 1 : 0;
    for (int s = 0; s < SHX_NUM_KERNELS; s++) { c->kt_ms[s] = 0.0; c->kt_n[s] = 0; }
    return SHX_OK;
}

extern "C" int shx_kernel_times(shx_case *c, const char *names[SHX_NUM_KERNELS], double ms[SHX_NUM_KERNELS],
                                int launches[SHX_NUM_KERNELS])
{
    if (!c) return set_err(SHX_ERR_ARG, "null handle");
    int rc = drain_events(c);
    if (rc) return rc;
    for (int s = 0; s < SHX_NUM_KERNELS; s++) {
        if (names) names[s] = g_knames[s];
        if (ms) ms[s] = c->kt_ms[s];
        if (launches) launches[s] = c->kt_n[s];
    }
    return SHX_OK;
}
