/**
 * bun:ffi drop-in for the reference's src/core energy / gradient / Sobolev-step
 * API, backed by librcurves.so (include/rcurves.h). Same signatures as
 * src/core/tangentPointEnergy.ts (calculateEnergy, gradientAnalytical,
 * gradientFiniteDiff) and src/core/optimizer.ts (step, sobolevStepSet); Vec3[] / Edge[] are flattened to
 * Float64Array / Int32Array at the boundary. One rc_ctx per topology, cached
 * like the worker's topology cache (src/worker/solverWorker.ts:32-43).
 *
 * Not executed in this repository's tests (no Bun runtime in the build
 * image); the Python host mirror (rcurves/) exercises the same C ABI.
 */
import { dlopen, FFIType, ptr, type Pointer } from 'bun:ffi';

type Vec3 = [number, number, number];
type Edge = [number, number];

const LIB = process.env.RCURVES_LIB ?? new URL('../librcurves.so', import.meta.url).pathname;

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
    vertices: Vec3[], edges: Edge[], _disjointPairs: number[][],
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
    vertices: Vec3[], edges: Edge[], _disjointPairs: number[][],
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
    vertices: Vec3[], edges: Edge[], _disjointPairs: number[][],
    alpha: number, beta: number, epsilon: number, h: number,
): Vec3[] {
    const c = contextFor(edges, vertices.length);
    const g = new Float64Array(vertices.length * 3);
    const st = rc.rc_gradient_fd(c, ptr(flatVerts(vertices)), ptr(params(alpha, beta, epsilon)), h, ptr(g));
    if (st !== 0) throw new Error(`rc_gradient_fd: ${rc.rc_last_error(c)}`);
    return unflat(g);
}

export interface StepOptions {
    mode: 'analytical' | 'finiteDiff';
    stepSize: number;
    alpha?: number; beta?: number; epsilon?: number; h?: number;
    collectField?: boolean;
}

/** step (optimizer.ts:40-68): the raw L2 step. Pure: new arrays. */
export function step(
    vertices: Vec3[], edges: Edge[], _disjointPairs: number[][], opts: StepOptions,
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

/** POD block for the ConstraintSet builders (constraintSet.ts:87-301). */
export type PodBlock =
    | { kind: 'barycenter'; x0: Vec3 }
    | { kind: 'totalLength'; L0: number }
    | { kind: 'edgeLengths'; ell0: number[] }
    | { kind: 'point'; vertexIndex: number; target: Vec3 };

const BLOCK_BYTES = 80; // rc_block: 2 x i32, 7 x f64, ptr, 2 x i32
function packBlocks(set: PodBlock[]): { buf: ArrayBuffer; keep: Float64Array[] } {
    const buf = new ArrayBuffer(Math.max(1, set.length) * BLOCK_BYTES);
    const dv = new DataView(buf);
    const keep: Float64Array[] = [];
    set.forEach((b, i) => {
        const o = i * BLOCK_BYTES;
        const kind = { barycenter: 0, totalLength: 1, edgeLengths: 2, point: 3 }[b.kind];
        dv.setInt32(o, kind, true);
        if (b.kind === 'point') dv.setInt32(o + 4, b.vertexIndex, true);
        if (b.kind === 'barycenter') b.x0.forEach((x, c) => dv.setFloat64(o + 8 + 8 * c, x, true));
        if (b.kind === 'totalLength') dv.setFloat64(o + 32, b.L0, true);
        if (b.kind === 'point') b.target.forEach((x, c) => dv.setFloat64(o + 40 + 8 * c, x, true));
        if (b.kind === 'edgeLengths') {
            const arr = Float64Array.from(b.ell0);
            keep.push(arr);
            dv.setBigUint64(o + 64, BigInt(ptr(arr)), true);
            dv.setInt32(o + 72, arr.length, true);
        }
    });
    return { buf, keep };
}

export interface SobolevStepOptions {
    mode: 'analytical' | 'finiteDiff';
    alpha?: number; beta?: number; epsilon?: number; h?: number;
    energyBefore?: number;
    penalties?: { totalLength?: number; lengthDiff?: number; field?: { weight: number; X: Vec3 } };
    projectionMode?: 'frozen' | 'reassemble';
    collectTimings?: boolean;
    collectField?: boolean;
    precision?: 'fp64' | 'mixed';
}

/** sobolevStepSet (optimizer.ts:231-440) over the device path. */
export function sobolevStepSet(
    vertices: Vec3[], edges: Edge[], _disjointPairs: number[][], set: PodBlock[],
    opts: SobolevStepOptions,
) {
    const h = contextFor(edges, vertices.length);
    const o = new ArrayBuffer(128); // sizeof(rc_step_opts)
    rc.rc_default_step_opts(ptr(o));
    const dv = new DataView(o);
    dv.setFloat64(0, opts.alpha ?? 3, true);
    dv.setFloat64(8, opts.beta ?? 6, true);
    dv.setFloat64(16, opts.epsilon ?? 1e-10, true);
    if (opts.energyBefore !== undefined) {
        dv.setInt32(24, 1, true);
        dv.setFloat64(32, opts.energyBefore, true);
    }
    dv.setInt32(28, opts.precision === 'mixed' ? 1 : 0, true);
    dv.setInt32(40, opts.projectionMode === 'frozen' ? 1 : 0, true); // default 'reassemble' (optimizer.ts:144-160)
    dv.setInt32(108, opts.collectTimings ? 1 : 0, true);
    dv.setInt32(112, opts.mode === 'finiteDiff' ? 1 : 0, true);
    dv.setFloat64(120, opts.h ?? 1e-6, true);
    const pen = new Float64Array([
        opts.penalties?.totalLength ?? 0, opts.penalties?.lengthDiff ?? 0,
        opts.penalties?.field?.weight ?? 0, ...(opts.penalties?.field?.X ?? [0, 0, 0]),
    ]);
    const { buf, keep } = packBlocks(set);
    const out = new Float64Array(vertices.length * 3);
    const stats = new ArrayBuffer(88);
    const field = opts.collectField ? new Float64Array(vertices.length * 3) : null;
    const timings = opts.collectTimings ? new Float64Array(8) : null;
    const st = rc.rc_sobolev_step(
        h, ptr(flatVerts(vertices)), ptr(buf), set.length, ptr(pen), ptr(o), ptr(out), ptr(stats),
        field ? ptr(field) : null, timings ? ptr(timings) : null,
    );
    void keep;
    if (st !== 0) throw new Error(`rc_sobolev_step: ${rc.rc_last_error(h)}`);
    const s = new DataView(stats);
    const projIt = s.getInt32(72, true);
    const reason = REASONS[s.getInt32(76, true)];
    const keys = ['dE', 'energy', 'assembleA', 'saddle', 'factor', 'projection', 'lineSearch', 'step'];
    return {
        vertices: unflat(out),
        energy: s.getFloat64(0, true),
        accepted: s.getInt32(64, true) === 1,
        converged: s.getInt32(68, true) === 1,
        stats: {
            tau: s.getFloat64(16, true),
            residual: s.getFloat64(24, true),
            gradientL2Norm: s.getFloat64(32, true),
            projectionIterations: projIt < 0 ? null : projIt,
            ...(reason ? { reason } : {}),
        },
        ...(timings ? { timings: Object.fromEntries(keys.map((k, i) => [k, { ms: timings[i], calls: 1 }])) } : {}),
        ...(field ? { descentField: reason === 'singular_system' ? null : unflat(field) } : {}),
    };
}
