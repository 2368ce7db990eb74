/**
 * bun:ffi drop-in for the reference's src/core energy / gradient / Sobolev-step
 * API, backed by librcurves.so (include/rcurves.h). Same signatures as
 * src/core/tangentPointEnergy.ts (calculateEnergy, gradientAnalytical,
 * gradientFiniteDiff), src/core/optimizer.ts (step, sobolevStep,
 * sobolevStepSet), src/core/sobolev/constraintSet.ts (the four block builders,
 * evaluateConstraintSet, assertValidConstraintSet), src/core/arrowField.ts
 * (computeArrowField), src/core/dispatch.ts (dispatchDescentStep) and the
 * message handler of src/worker/solverWorker.ts. Vec3[] / Edge[] are
 * flattened to Float64Array / Int32Array at the boundary. One rc_ctx per
 * topology, cached like the worker's topology cache (solverWorker.ts:32-43).
 *
 * The block builders return the reference's ConstraintBlock (kind, evaluate,
 * scale) with the block's frozen target beside the closures: the device
 * evaluates constraints from that POD form (rc_block), the closures serve
 * callers that evaluate a block on the host. A block without the POD fields
 * (hand-written closures) has no device form and is rejected.
 *
 * Struct sizes and field offsets below follow include/rcurves.h; the CPU test
 * suite compares them with a compiled probe of the header. Not executed in
 * this repository's tests (no Bun runtime in the build image); the Python
 * host mirror (rcurves/) exercises the same C ABI.
 */
import { dlopen, FFIType, ptr, type Pointer } from 'bun:ffi';

type Vec3 = [number, number, number];
type Edge = [number, number];

const LIB = process.env.RCURVES_LIB ?? new URL('../librcurves.so', import.meta.url).pathname;

// ---- C struct layouts (include/rcurves.h) ----------------------------------
const OPTS_BYTES = 152; // rc_step_opts
const STATS_BYTES = 112; // rc_step_stats
const TIMINGS_BYTES = 128; // rc_timings: 8 x rc_phase {f64 ms, i32 calls, i32 pad}
const PHASE_BYTES = 16; // rc_phase
const BLOCK_BYTES = 80; // rc_block
const OPT = {
    alpha: 0, beta: 8, epsilon: 16, has_energy_before: 24, precision: 28, energy_before: 32,
    projection_frozen: 40, collect_timings: 108, gradient_mode: 112, fd_h: 120,
} as const;
const STAT = {
    energy: 0, tau: 16, residual: 24, gradient_l2_norm: 32, accepted: 64, converged: 68,
    projection_iterations: 72, reason: 76,
} as const;
const BLK = { kind: 0, vertex: 4, x0: 8, L0: 32, target: 40, ell0: 64, ell0_len: 72 } as const;
const PHASE_KEYS = [
    'dE', 'energy', 'assembleA', 'saddle', 'factor', 'projection', 'lineSearch', 'step',
] as const;

const { symbols: rc } = dlopen(LIB, {
    rc_ctx_create: { args: [FFIType.ptr, FFIType.i32, FFIType.i32, FFIType.ptr], returns: FFIType.i32 },
    rc_ctx_destroy: { args: [FFIType.ptr], returns: FFIType.void },
    rc_last_error: { args: [FFIType.ptr], returns: FFIType.cstring },
    rc_energy: { args: [FFIType.ptr, FFIType.ptr, FFIType.ptr, FFIType.ptr], returns: FFIType.i32 },
    rc_gradient: {
        args: [FFIType.ptr, FFIType.ptr, FFIType.ptr, FFIType.ptr, FFIType.ptr],
        returns: FFIType.i32,
    },
    rc_gradient_fd: {
        args: [FFIType.ptr, FFIType.ptr, FFIType.ptr, FFIType.f64, FFIType.ptr],
        returns: FFIType.i32,
    },
    rc_step: {
        args: [
            FFIType.ptr, FFIType.ptr, FFIType.ptr, FFIType.i32, FFIType.f64, FFIType.f64,
            FFIType.ptr, FFIType.ptr, FFIType.ptr,
        ],
        returns: FFIType.i32,
    },
    rc_default_step_opts: { args: [FFIType.ptr], returns: FFIType.void },
    rc_constraint_rows: { args: [FFIType.ptr, FFIType.ptr, FFIType.i32], returns: FFIType.i32 },
    rc_solve_constrained_gradient: {
        args: [
            FFIType.ptr, FFIType.ptr, FFIType.ptr, FFIType.ptr, FFIType.i32, FFIType.ptr,
            FFIType.ptr, FFIType.ptr, FFIType.ptr,
        ],
        returns: FFIType.i32,
    },
    rc_sobolev_step: {
        args: [
            FFIType.ptr, FFIType.ptr, FFIType.ptr, FFIType.i32, FFIType.ptr,
            FFIType.ptr, FFIType.ptr, FFIType.ptr, FFIType.ptr, FFIType.ptr,
        ],
        returns: FFIType.i32,
    },
});

const REASONS = [
    undefined, 'zero_or_nonfinite_gradient_norm', 'not_a_descent_direction', 'armijo_failed',
    'projection_failed', 'tau_below_min', 'singular_system',
] as const;

// ---- topology cache -------------------------------------------------------
const ctxCache = new Map<string, Pointer>();
function contextFor(edges: Edge[], nVertices: number): Pointer {
    const key = `${nVertices}:${edges.length}:${edges.map((e) => `${e[0]},${e[1]}`).join(';')}`;
    const hit = ctxCache.get(key);
    if (hit) return hit;
    const flat = new Int32Array(edges.length * 2);
    edges.forEach(([a, b], i) => { flat[2 * i] = a; flat[2 * i + 1] = b; });
    const out = new BigUint64Array(1);
    const st = rc.rc_ctx_create(ptr(flat), edges.length, nVertices, ptr(out));
    const handle = Number(out[0]) as unknown as Pointer;
    if (st !== 0) {
        const msg = handle ? String(rc.rc_last_error(handle)) : 'rc_ctx_create failed';
        if (handle) rc.rc_ctx_destroy(handle);
        throw new Error(`rc_ctx_create: ${msg}`);
    }
    ctxCache.set(key, handle);
    return handle;
}

function flatVerts(v: Vec3[]): Float64Array {
    const f = new Float64Array(v.length * 3);
    v.forEach((p, i) => { f[3 * i] = p[0]; f[3 * i + 1] = p[1]; f[3 * i + 2] = p[2]; });
    return f;
}
function unflat(f: Float64Array): Vec3[] {
    const out: Vec3[] = [];
    for (let i = 0; i < f.length; i += 3) out.push([f[i], f[i + 1], f[i + 2]]);
    return out;
}
function params(alpha: number, beta: number, epsilon: number): Float64Array {
    return new Float64Array([alpha, beta, epsilon]);
}

/** calculateEnergy (tangentPointEnergy.ts:50-111). disjointPairs is accepted for signature parity. */
export function calculateEnergy(
    vertices: Vec3[], edges: Edge[], _disjointPairs: number[][] | null,
    alpha: number, beta: number, epsilon: number,
): number {
    const h = contextFor(edges, vertices.length);
    const out = new Float64Array(1);
    const st = rc.rc_energy(h, ptr(flatVerts(vertices)), ptr(params(alpha, beta, epsilon)), ptr(out));
    if (st !== 0) throw new Error(`rc_energy: ${rc.rc_last_error(h)}`);
    return out[0];
}

/** gradientAnalytical (tangentPointEnergy.ts:139-392). */
export function gradientAnalytical(
    vertices: Vec3[], edges: Edge[], _disjointPairs: number[][] | null,
    alpha: number, beta: number, epsilon: number,
): Vec3[] {
    const h = contextFor(edges, vertices.length);
    const g = new Float64Array(vertices.length * 3);
    const st = rc.rc_gradient(h, ptr(flatVerts(vertices)), ptr(params(alpha, beta, epsilon)), ptr(g), null);
    if (st !== 0) throw new Error(`rc_gradient: ${rc.rc_last_error(h)}`);
    return unflat(g);
}

/** gradientFiniteDiff (tangentPointEnergy.ts:114-136). */
export function gradientFiniteDiff(
    vertices: Vec3[], edges: Edge[], _disjointPairs: number[][] | null,
    alpha: number, beta: number, epsilon: number, h: number,
): Vec3[] {
    const c = contextFor(edges, vertices.length);
    const g = new Float64Array(vertices.length * 3);
    const st = rc.rc_gradient_fd(c, ptr(flatVerts(vertices)), ptr(params(alpha, beta, epsilon)), h, ptr(g));
    if (st !== 0) throw new Error(`rc_gradient_fd: ${rc.rc_last_error(c)}`);
    return unflat(g);
}

/**
 * Function-style forward/backward pair over the device kernels: forward()
 * evaluates the tangent-point energy (rc_energy), backward() the adjoint
 * gradient of the same energy at the saved vertices (rc_gradient, which
 * recomputes the pair terms instead of keeping an E^2 tape), scaled by the
 * upstream gradient. The curve is kept between the two calls, as an autograd
 * Function keeps its saved tensors.
 */
export class TangentPointEnergy {
    private saved: Vec3[] | null = null;

    constructor(
        readonly edges: Edge[], readonly numVertices: number,
        readonly alpha = 3, readonly beta = 6, readonly epsilon = 1e-10,
    ) {}

    forward(vertices: Vec3[]): number {
        if (vertices.length !== this.numVertices) throw new Error('vertex count differs from the topology');
        this.saved = vertices.map((p) => [p[0], p[1], p[2]] as Vec3);
        return calculateEnergy(this.saved, this.edges, null, this.alpha, this.beta, this.epsilon);
    }

    backward(gradOutput = 1): Vec3[] {
        if (!this.saved) throw new Error('backward() before forward()');
        const g = gradientAnalytical(this.saved, this.edges, null, this.alpha, this.beta, this.epsilon);
        return g.map((p) => [gradOutput * p[0], gradOutput * p[1], gradOutput * p[2]] as Vec3);
    }
}

export interface StepOptions {
    mode: 'analytical' | 'finiteDiff';
    stepSize: number;
    alpha?: number; beta?: number; epsilon?: number; h?: number;
    collectField?: boolean;
}

/** step (optimizer.ts:40-68): the raw L2 step. Pure: new arrays. */
export function step(
    vertices: Vec3[], edges: Edge[], _disjointPairs: number[][] | null, opts: StepOptions,
): { vertices: Vec3[]; energy: number; field?: Vec3[] } {
    const c = contextFor(edges, vertices.length);
    const out = new Float64Array(vertices.length * 3);
    const energy = new Float64Array(1);
    const field = opts.collectField ? new Float64Array(vertices.length * 3) : null;
    const st = rc.rc_step(
        c, ptr(flatVerts(vertices)), ptr(params(opts.alpha ?? 3, opts.beta ?? 6, opts.epsilon ?? 1e-10)),
        opts.mode === 'analytical' ? 0 : 1, opts.stepSize, opts.h ?? 1e-6, ptr(out), ptr(energy),
        field ? ptr(field) : null,
    );
    if (st !== 0) throw new Error(`rc_step: ${rc.rc_last_error(c)}`);
    if (field) return { vertices: unflat(out), energy: energy[0], field: unflat(field) };
    return { vertices: unflat(out), energy: energy[0] };
}

// ---- constraint blocks (constraintSet.ts:26-344) ---------------------------

export interface ConstraintEval {
    phi: number[];
    C: number[][];
}

/** The frozen target of a block: what the device needs (rc_block). */
export type PodBlock =
    | { kind: 'barycenter'; x0: Vec3 }
    | { kind: 'totalLength'; L0: number }
    | { kind: 'edgeLengths'; ell0: number[] }
    | { kind: 'point'; vertexIndex: number; target: Vec3 };

/** ConstraintBlock (constraintSet.ts:39-50) carrying its POD target. */
export type ConstraintBlock = PodBlock & {
    evaluate(vertices: Vec3[], edges: Edge[]): ConstraintEval;
    scale(vertices: Vec3[], edges: Edge[]): number;
};
export type ConstraintSet = ConstraintBlock[];

// Row index of coordinate c of vertex i in a 3n row (layout.ts:56-58).
const col = (c: number, i: number, n: number): number => c * n + i;

// Edge vector, raw length and unit tangent, T = 0 below the 1e-14 guard
// (constraintSet.ts:135-141, constraints.ts:89-98).
function edgeFrame(p1: Vec3, p2: Vec3): { ell: number; T: Vec3 } {
    const d: Vec3 = [p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]];
    const ell = Math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (ell < 1e-14) return { ell, T: [0, 0, 0] };
    return { ell, T: [d[0] / ell, d[1] / ell, d[2] / ell] };
}

/** totalLength (constraintSet.ts:64-76). */
export function totalLength(vertices: Vec3[], edges: Edge[]): number {
    let L = 0;
    for (const [a, b] of edges) L += edgeFrame(vertices[a], vertices[b]).ell;
    return L;
}

/** edgeLengths (constraintSet.ts:172-181). */
export function edgeLengths(vertices: Vec3[], edges: Edge[]): number[] {
    return edges.map(([a, b]) => edgeFrame(vertices[a], vertices[b]).ell);
}

/** barycenterBlock (constraintSet.ts:87-99): Phi = sum_I l_I (m_I - x0),
 * scale max(1, L max(1, R)) with R the largest vertex distance from x0
 * (lineSearch.ts:86-109). */
export function barycenterBlock(x0: Vec3): ConstraintBlock {
    return {
        kind: 'barycenter',
        x0,
        evaluate(vertices, edges) {
            const n = vertices.length;
            const phi = [0, 0, 0];
            const C = [0, 1, 2].map(() => new Array<number>(3 * n).fill(0));
            for (const [a, b] of edges) {
                const p1 = vertices[a], p2 = vertices[b];
                const { ell, T } = edgeFrame(p1, p2);
                const r = [0, 1, 2].map((c) => 0.5 * (p1[c] + p2[c]) - x0[c]);
                for (let i = 0; i < 3; i++) {
                    phi[i] += ell * r[i];
                    for (let c = 0; c < 3; c++) {
                        const half = i === c ? 0.5 * ell : 0;
                        C[i][col(c, a, n)] += -T[c] * r[i] + half;
                        C[i][col(c, b, n)] += T[c] * r[i] + half;
                    }
                }
            }
            return { phi, C };
        },
        scale(vertices, edges) {
            let R = 0;
            for (const [a, b] of edges)
                for (const p of [vertices[a], vertices[b]])
                    R = Math.max(R, Math.hypot(p[0] - x0[0], p[1] - x0[1], p[2] - x0[2]));
            return Math.max(1, totalLength(vertices, edges) * Math.max(1, R));
        },
    };
}

/** totalLengthBlock (constraintSet.ts:119-157): Phi = L0 - L. */
export function totalLengthBlock(L0: number): ConstraintBlock {
    return {
        kind: 'totalLength',
        L0,
        evaluate(vertices, edges) {
            const n = vertices.length;
            const row = new Array<number>(3 * n).fill(0);
            let L = 0;
            for (const [a, b] of edges) {
                const { ell, T } = edgeFrame(vertices[a], vertices[b]);
                L += ell;
                for (let c = 0; c < 3; c++) {
                    row[col(c, a, n)] += T[c];
                    row[col(c, b, n)] -= T[c];
                }
            }
            return { phi: [L0 - L], C: [row] };
        },
        scale: (vertices, edges) => Math.max(1, totalLength(vertices, edges)),
    };
}

/** edgeLengthsBlock (constraintSet.ts:206-258): Phi_I = l0_I - l_I, NaN where
 * the target vector is short. */
export function edgeLengthsBlock(ell0: number[]): ConstraintBlock {
    return {
        kind: 'edgeLengths',
        ell0,
        evaluate(vertices, edges) {
            const n = vertices.length;
            const phi: number[] = [];
            const C: number[][] = [];
            edges.forEach(([a, b], r) => {
                const { ell, T } = edgeFrame(vertices[a], vertices[b]);
                phi.push((ell0[r] ?? Number.NaN) - ell);
                const row = new Array<number>(3 * n).fill(0);
                for (let c = 0; c < 3; c++) {
                    row[col(c, a, n)] += T[c];
                    row[col(c, b, n)] -= T[c];
                }
                C.push(row);
            });
            return { phi, C };
        },
        scale: (vertices, edges) => Math.max(1, totalLength(vertices, edges)),
    };
}

/** pointBlock (constraintSet.ts:271-301): Phi = gamma_i - target; an
 * out-of-range index gives NaN Phi and zero rows. */
export function pointBlock(vertexIndex: number, target: Vec3): ConstraintBlock {
    return {
        kind: 'point',
        vertexIndex,
        target,
        evaluate(vertices, _edges) {
            const n = vertices.length;
            const C = [0, 1, 2].map(() => new Array<number>(3 * n).fill(0));
            const p = vertices[vertexIndex];
            if (p === undefined) return { phi: [Number.NaN, Number.NaN, Number.NaN], C };
            for (let r = 0; r < 3; r++) C[r][col(r, vertexIndex, n)] = 1;
            return { phi: [p[0] - target[0], p[1] - target[1], p[2] - target[2]], C };
        },
        scale(vertices, _edges) {
            let R = 0;
            for (const v of vertices)
                R = Math.max(R, Math.hypot(v[0] - target[0], v[1] - target[1], v[2] - target[2]));
            return Math.max(1, R);
        },
    };
}

/** evaluateConstraintSet (constraintSet.ts:309-325): blocks stacked in order. */
export function evaluateConstraintSet(set: ConstraintSet, vertices: Vec3[], edges: Edge[]): ConstraintEval {
    const phi: number[] = [];
    const C: number[][] = [];
    for (const b of set) {
        const r = b.evaluate(vertices, edges);
        phi.push(...r.phi);
        C.push(...r.C);
    }
    return { phi, C };
}

/** assertValidConstraintSet (constraintSet.ts:333-344). */
export function assertValidConstraintSet(set: ConstraintSet): void {
    if (set.some((b) => b.kind === 'totalLength') && set.some((b) => b.kind === 'edgeLengths'))
        throw new Error(
            'ConstraintSet: totalLength and edgeLengths are mutually exclusive — the total-length ' +
                'row is the sum of the edge-length rows (rank deficiency, spec §3.4)',
        );
}

function packBlocks(set: (PodBlock | ConstraintBlock)[]): { buf: ArrayBuffer; keep: Float64Array[] } {
    const buf = new ArrayBuffer(Math.max(1, set.length) * BLOCK_BYTES);
    const dv = new DataView(buf);
    const keep: Float64Array[] = [];
    set.forEach((b, i) => {
        const o = i * BLOCK_BYTES;
        switch (b.kind) {
            case 'barycenter':
                if (!('x0' in b)) break;
                dv.setInt32(o + BLK.kind, 0, true);
                b.x0.forEach((x, c) => dv.setFloat64(o + BLK.x0 + 8 * c, x, true));
                return;
            case 'totalLength':
                if (!('L0' in b)) break;
                dv.setInt32(o + BLK.kind, 1, true);
                dv.setFloat64(o + BLK.L0, b.L0, true);
                return;
            case 'edgeLengths': {
                if (!('ell0' in b)) break;
                const arr = Float64Array.from(b.ell0);
                keep.push(arr);
                dv.setInt32(o + BLK.kind, 2, true);
                if (arr.length > 0) dv.setBigUint64(o + BLK.ell0, BigInt(ptr(arr)), true);
                dv.setInt32(o + BLK.ell0_len, arr.length, true);
                return;
            }
            case 'point':
                if (!('vertexIndex' in b)) break;
                dv.setInt32(o + BLK.kind, 3, true);
                dv.setInt32(o + BLK.vertex, b.vertexIndex, true);
                b.target.forEach((x, c) => dv.setFloat64(o + BLK.target + 8 * c, x, true));
                return;
        }
        throw new Error(`constraint block ${i} (${(b as { kind: string }).kind}) has no device form`);
    });
    return { buf, keep };
}

// ---- the Sobolev step (optimizer.ts:95-469) --------------------------------

export interface PenaltyConfig {
    totalLength?: number; lengthDiff?: number; field?: { weight: number; X: Vec3 };
}
export interface SobolevStepOptions {
    mode: 'analytical' | 'finiteDiff';
    alpha?: number; beta?: number; epsilon?: number; h?: number;
    energyBefore?: number;
    penalties?: PenaltyConfig;
    projectionMode?: 'frozen' | 'reassemble';
    collectTimings?: boolean;
    collectField?: boolean;
    precision?: 'fp64' | 'mixed';
}
export type PhaseKey = (typeof PHASE_KEYS)[number];
export type SobolevStepTimings = Partial<Record<PhaseKey, { ms: number; calls: number }>>;
export interface SobolevStepStats {
    tau: number; residual: number; gradientL2Norm: number;
    projectionIterations: number | null; reason?: (typeof REASONS)[number];
}
export interface SobolevStepResult {
    vertices: Vec3[]; energy: number; accepted: boolean; converged: boolean;
    stats: SobolevStepStats; timings?: SobolevStepTimings; descentField?: Vec3[] | null;
}

function stepOpts(opts: SobolevStepOptions): ArrayBuffer {
    const o = new ArrayBuffer(OPTS_BYTES);
    rc.rc_default_step_opts(ptr(o));
    const dv = new DataView(o);
    dv.setFloat64(OPT.alpha, opts.alpha ?? 3, true);
    dv.setFloat64(OPT.beta, opts.beta ?? 6, true);
    dv.setFloat64(OPT.epsilon, opts.epsilon ?? 1e-10, true);
    if (opts.energyBefore !== undefined) {
        dv.setInt32(OPT.has_energy_before, 1, true);
        dv.setFloat64(OPT.energy_before, opts.energyBefore, true);
    }
    dv.setInt32(OPT.precision, opts.precision === 'mixed' ? 1 : 0, true);
    // default 'reassemble' (optimizer.ts:144-160)
    dv.setInt32(OPT.projection_frozen, opts.projectionMode === 'frozen' ? 1 : 0, true);
    dv.setInt32(OPT.collect_timings, opts.collectTimings ? 1 : 0, true);
    dv.setInt32(OPT.gradient_mode, opts.mode === 'finiteDiff' ? 1 : 0, true);
    dv.setFloat64(OPT.fd_h, opts.h ?? 1e-6, true);
    return o;
}

/** sobolevStepSet (optimizer.ts:231-440) over the device path. */
export function sobolevStepSet(
    vertices: Vec3[], edges: Edge[], _disjointPairs: number[][] | null,
    set: (PodBlock | ConstraintBlock)[], opts: SobolevStepOptions,
): SobolevStepResult {
    const h = contextFor(edges, vertices.length);
    const o = stepOpts(opts);
    const pen = new Float64Array([
        opts.penalties?.totalLength ?? 0, opts.penalties?.lengthDiff ?? 0,
        opts.penalties?.field?.weight ?? 0, ...(opts.penalties?.field?.X ?? [0, 0, 0]),
    ]);
    const { buf, keep } = packBlocks(set);
    const out = new Float64Array(vertices.length * 3);
    const stats = new ArrayBuffer(STATS_BYTES);
    const field = opts.collectField ? new Float64Array(vertices.length * 3) : null;
    const timings = opts.collectTimings ? new ArrayBuffer(TIMINGS_BYTES) : null;
    const st = rc.rc_sobolev_step(
        h, ptr(flatVerts(vertices)), ptr(buf), set.length, ptr(pen), ptr(o), ptr(out), ptr(stats),
        field ? ptr(field) : null, timings ? ptr(timings) : null,
    );
    void keep;
    if (st !== 0) throw new Error(`rc_sobolev_step: ${rc.rc_last_error(h)}`);
    const s = new DataView(stats);
    const projIt = s.getInt32(STAT.projection_iterations, true);
    const reason = REASONS[s.getInt32(STAT.reason, true)];
    const result: SobolevStepResult = {
        vertices: unflat(out),
        energy: s.getFloat64(STAT.energy, true),
        accepted: s.getInt32(STAT.accepted, true) === 1,
        converged: s.getInt32(STAT.converged, true) === 1,
        stats: {
            tau: s.getFloat64(STAT.tau, true),
            residual: s.getFloat64(STAT.residual, true),
            gradientL2Norm: s.getFloat64(STAT.gradient_l2_norm, true),
            projectionIterations: projIt < 0 ? null : projIt,
            ...(reason ? { reason } : {}),
        },
    };
    if (timings) {
        // only the phases that fired, like the reference's Partial ledger
        const t = new DataView(timings);
        const ledger: SobolevStepTimings = {};
        PHASE_KEYS.forEach((k, i) => {
            const calls = t.getInt32(i * PHASE_BYTES + 8, true);
            if (calls > 0) ledger[k] = { ms: t.getFloat64(i * PHASE_BYTES, true), calls };
        });
        result.timings = ledger;
    }
    if (field) result.descentField = reason === 'singular_system' ? null : unflat(field);
    return result;
}

/** sobolevStep (optimizer.ts:455-469): the barycenter-only set. */
export function sobolevStep(
    vertices: Vec3[], edges: Edge[], disjointPairs: number[][] | null, x0: Vec3,
    opts: SobolevStepOptions,
): SobolevStepResult {
    return sobolevStepSet(vertices, edges, disjointPairs, [barycenterBlock(x0)], opts);
}

// ---- arrows, dispatch and the worker (arrowField.ts, dispatch.ts, solverWorker.ts)

const DEFAULTS = { alpha: 3, beta: 6, epsilon: 1e-10, h: 1e-6 } as const;

/** computeArrowField (arrowField.ts:24-74): dE at the defaults; 'sobolev'
 * maps it to the barycenter-only g~ (null when the solve throws). */
export function computeArrowField(
    vertices: Vec3[], edges: Edge[], disjointPairs: number[][] | null,
    mode: 'analytical' | 'finiteDiff', descentMode: 'sobolev' | 'raw', x0: Vec3,
): Vec3[] | null {
    const { alpha, beta, epsilon, h } = DEFAULTS;
    const dE = mode === 'analytical'
        ? gradientAnalytical(vertices, edges, disjointPairs, alpha, beta, epsilon)
        : gradientFiniteDiff(vertices, edges, disjointPairs, alpha, beta, epsilon, h);
    if (descentMode !== 'sobolev') return dE;
    const c = contextFor(edges, vertices.length);
    const { buf } = packBlocks([barycenterBlock(x0)]);
    const k = rc.rc_constraint_rows(c, ptr(buf), 1);
    const gt = new Float64Array(vertices.length * 3);
    const lambda = new Float64Array(Math.max(1, k));
    const res = new Float64Array(1);
    const st = rc.rc_solve_constrained_gradient(
        c, ptr(flatVerts(vertices)), ptr(flatVerts(dE)), ptr(buf), 1, null, ptr(gt), ptr(lambda),
        ptr(res),
    );
    return st === 0 ? unflat(gt) : null;
}

export interface PinArg { vertexIndex: number; target: Vec3; enabled: boolean }
export interface DispatchDescentStepArgs {
    descentMode: 'sobolev' | 'raw';
    vertices: Vec3[]; edges: Edge[]; disjointPairs: number[][] | null;
    mode: 'analytical' | 'finiteDiff'; stepSize: number; x0: Vec3;
    barycenterConstraint?: boolean; lengthConstraint?: boolean;
    lengthMode?: 'total' | 'perEdge' | 'none';
    sobolevL0?: number; sobolevEll0?: number[]; pins?: PinArg[];
    collectTimings?: boolean; collectField?: boolean; energyBefore?: number;
    projectionMode?: 'frozen' | 'reassemble'; penalties?: PenaltyConfig;
    precision?: 'fp64' | 'mixed';
}
export interface DescentStepOutcome {
    vertices: Vec3[]; energy: number; accepted: boolean; converged: boolean;
    stats: SobolevStepStats | null; timings: SobolevStepTimings | null;
    descentField: Vec3[] | null;
}

/** dispatchDescentStep (dispatch.ts:185-290). */
export function dispatchDescentStep(args: DispatchDescentStepArgs): DescentStepOutcome {
    if (args.descentMode !== 'sobolev') {
        const r = step(args.vertices, args.edges, args.disjointPairs, {
            mode: args.mode, stepSize: args.stepSize, collectField: args.collectField,
        });
        return {
            vertices: r.vertices, energy: r.energy, accepted: true, converged: false,
            stats: null, timings: null, descentField: r.field ?? null,
        };
    }
    if (args.barycenterConstraint === undefined && args.lengthConstraint === undefined &&
        args.lengthMode === undefined) {
        // pre-toggle call shape: barycenter-only, the mode alone, nothing collected
        const r = sobolevStep(args.vertices, args.edges, args.disjointPairs, args.x0, { mode: args.mode });
        return {
            vertices: r.vertices, energy: r.energy, accepted: r.accepted, converged: r.converged,
            stats: r.stats, timings: null, descentField: null,
        };
    }
    const set: ConstraintSet = [];
    if (args.barycenterConstraint ?? true) set.push(barycenterBlock(args.x0));
    const lengthMode = args.lengthMode ?? ((args.lengthConstraint ?? true) ? 'total' : 'none');
    if (lengthMode === 'total') set.push(totalLengthBlock(args.sobolevL0 ?? Number.NaN));
    else if (lengthMode === 'perEdge')
        set.push(edgeLengthsBlock(args.sobolevEll0 ?? args.edges.map(() => Number.NaN)));
    for (const pin of args.pins ?? [])
        if (pin.enabled && pin.vertexIndex >= 0 && pin.vertexIndex < args.vertices.length)
            set.push(pointBlock(pin.vertexIndex, pin.target));
    assertValidConstraintSet(set);
    const r = sobolevStepSet(args.vertices, args.edges, args.disjointPairs, set, {
        mode: args.mode, collectTimings: args.collectTimings, collectField: args.collectField,
        energyBefore: args.energyBefore, projectionMode: args.projectionMode,
        penalties: args.penalties, precision: args.precision,
    });
    return {
        vertices: r.vertices, energy: r.energy, accepted: r.accepted, converged: r.converged,
        stats: r.stats, timings: r.timings ?? null, descentField: r.descentField ?? null,
    };
}

export type SolverWorkerRequest =
    | { type: 'topology'; graphVersion: number; edges: Edge[] }
    | {
          type: 'field'; graphVersion: number;
          args: { vertices: Vec3[]; mode: 'analytical' | 'finiteDiff'; descentMode: 'sobolev' | 'raw'; x0: Vec3 };
      }
    | { type: 'step'; graphVersion: number; args: Omit<DispatchDescentStepArgs, 'edges' | 'disjointPairs'> };
export type SolverWorkerResponse =
    | { type: 'result'; graphVersion: number; result: DescentStepOutcome }
    | { type: 'fieldResult'; graphVersion: number; field: Vec3[] | null }
    | { type: 'error'; message: string };

/**
 * The solver worker's message handler (solverWorker.ts:34-98) over the device
 * path: a topology cache per graphVersion (the device context per topology is
 * cached by contextFor, the disjoint pairs are derived on the device), step
 * and field requests before any topology are errors, graphVersion is echoed
 * untouched, and any throw becomes an {type:'error'} response. Returns the
 * response (null for a topology message, which has none).
 */
export function createSolverWorkerHandler(): (msg: SolverWorkerRequest) => SolverWorkerResponse | null {
    let topology: { graphVersion: number; edges: Edge[] } | null = null;
    return (msg) => {
        try {
            if (msg.type === 'topology') {
                topology = { graphVersion: msg.graphVersion, edges: msg.edges };
                return null;
            }
            if (msg.type === 'field') {
                if (topology === null) throw new Error('solverWorker: field received before topology');
                const a = msg.args;
                const field = computeArrowField(a.vertices, topology.edges, null, a.mode, a.descentMode, a.x0);
                return { type: 'fieldResult', graphVersion: msg.graphVersion, field };
            }
            if (topology === null) throw new Error('solverWorker: step received before topology');
            const result = dispatchDescentStep({ ...msg.args, edges: topology.edges, disjointPairs: null });
            return { type: 'result', graphVersion: msg.graphVersion, result };
        } catch (err) {
            return { type: 'error', message: err instanceof Error ? err.message : String(err) };
        }
    };
}

/** Installs the handler on a worker scope (`self` inside a Bun/browser worker). */
export function installSolverWorker(scope: {
    onmessage: ((ev: MessageEvent<SolverWorkerRequest>) => void) | null;
    postMessage(msg: SolverWorkerResponse): void;
}): void {
    const handle = createSolverWorkerHandler();
    scope.onmessage = (ev) => {
        const r = handle(ev.data);
        if (r !== null) scope.postMessage(r);
    };
}
