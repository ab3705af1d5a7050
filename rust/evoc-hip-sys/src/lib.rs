//! Raw bindings to `include/evoc_hip.h`. Every function is `extern "C"`, never unwinds,
//! and returns an `i32` status (`EVOC_OK` = 0); `evoc_hip_last_error` holds the calling
//! thread's message. Buffers are borrowed for the duration of a call; results are owned by
//! the library and released with `evoc_hip_result_free`.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_double, c_float, c_int};

pub const EVOC_OK: c_int = 0;
pub const EVOC_ERR_INVALID_ARGUMENT: c_int = 1;
pub const EVOC_ERR_HIP: c_int = 2;
pub const EVOC_ERR_NO_DEVICE: c_int = 3;
pub const EVOC_ERR_INTERNAL: c_int = 4;

pub const EVOC_METRIC_EUCLIDEAN: i32 = 0;
pub const EVOC_METRIC_COSINE: i32 = 1;

pub const EVOC_KNN_AUTO: i32 = 0;
pub const EVOC_KNN_SPLIT: i32 = 1;
pub const EVOC_KNN_BF16: i32 = 2;
pub const EVOC_KNN_FP8: i32 = 3;

/// Mirrors `EvocParams` (src/lib.rs:52-80 of evoc-rs).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct evoc_params_t {
    pub n_neighbours: i32,
    pub noise_level: c_float,
    pub n_epochs: i32,
    pub embedding_dim: i32,
    pub neighbour_scale: c_float,
    pub symmetrise: i32,
    pub min_samples: i32,
    pub base_min_cluster_size: i32,
    pub approx_n_clusters: i32,
    pub min_similarity_threshold: c_double,
    pub max_layers: i32,
}

/// `dist_metric` of `NearestNeighbourParamsEvoc` plus the device settings.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct evoc_hip_nn_params_t {
    pub metric: i32,
    pub knn_precision: i32,
    pub warm_start: i32,
    pub warm_epochs: i32,
    pub mlp_subsample: i32,
    pub device_id: i32,
}

/// Mirrors `EvocResult` (src/lib.rs:106-118 of evoc-rs); owned by the library.
#[repr(C)]
#[derive(Debug)]
pub struct evoc_result_t {
    pub n: i64,
    pub k: i32,
    pub n_layers: i32,
    pub dim: i32,
    pub best_layer: i32,
    pub cluster_layers: *mut i64,
    pub membership_strengths: *mut c_float,
    pub persistence_scores: *mut c_double,
    pub nn_indices: *mut i32,
    pub nn_distances: *mut c_float,
    pub embedding: *mut c_float,
    pub stage_ms: [c_double; 8],
    pub knn_fallback_rows: i64,
    pub knn_kernel_ms: c_double,
    pub embed_kernel_ms: c_double,
    pub warm_start_used: i32,
    pub warm_start_ari: c_double,
    pub warm_start_mst_ratio: c_double,
}

extern "C" {
    pub fn evoc_hip_default_params(p: *mut evoc_params_t);
    pub fn evoc_hip_default_nn_params(p: *mut evoc_hip_nn_params_t);

    /// `evoc()` on a host matrix with arbitrary element strides.
    pub fn evoc_hip_run(
        data: *const c_float,
        n: i64,
        d: i64,
        row_stride: i64,
        col_stride: i64,
        pre_idx: *const i32,
        pre_dist: *const c_float,
        pre_k: i32,
        params: *const evoc_params_t,
        nn: *const evoc_hip_nn_params_t,
        seed: u64,
        verbose: i32,
        out: *mut *mut evoc_result_t,
    ) -> c_int;

    /// `evoc()` on a row-major matrix already in device memory.
    pub fn evoc_hip_run_device(
        d_data: *const c_float,
        n: i64,
        d: i64,
        ld: i64,
        params: *const evoc_params_t,
        nn: *const evoc_hip_nn_params_t,
        seed: u64,
        verbose: i32,
        out: *mut *mut evoc_result_t,
    ) -> c_int;

    pub fn evoc_hip_result_free(r: *mut evoc_result_t);
    pub fn evoc_hip_last_error() -> *const c_char;
    pub fn evoc_hip_device_count(count: *mut i32) -> c_int;
    pub fn evoc_hip_release() -> c_int;

    pub fn evoc_hip_knn(
        x: *const c_float,
        n: i64,
        d: i64,
        metric: i32,
        k: i32,
        precision: i32,
        idx_out: *mut i32,
        dist_out: *mut c_float,
        fallback_rows: *mut i64,
        kernel_ms: *mut c_double,
    ) -> c_int;

    pub fn evoc_hip_knn_last_stats(
        kernel_flop: *mut c_double,
        kernel_ms: *mut c_double,
        symmetric: *mut i32,
        escalated_rows: *mut i64,
    ) -> c_int;

    pub fn evoc_hip_knn_tile_pass(
        x: *const c_float,
        n: i64,
        d: i64,
        metric: i32,
        k: i32,
        precision: i32,
        cand_idx: *mut i32,
        cand_key: *mut c_float,
        key_scale: *mut c_float,
        ncand: *mut i32,
        gamma: *mut c_float,
    ) -> c_int;

    pub fn evoc_hip_fuzzy(
        knn_idx: *const i32,
        knn_dist: *const c_float,
        n: i64,
        k: i32,
        effective_k: c_float,
        symmetrise: i32,
        sigma: *mut c_float,
        rho: *mut c_float,
        indptr: *mut i64,
        indices: *mut i32,
        weights: *mut c_float,
        nnz: *mut i64,
    ) -> c_int;

    pub fn evoc_hip_embed(
        indptr: *const i64,
        indices: *const i32,
        weights: *const c_float,
        n: i64,
        dim: i32,
        n_epochs: i32,
        noise_level: c_float,
        negative_sample_rate: c_float,
        initial_alpha: c_float,
        beta1: c_float,
        beta2: c_float,
        eps: c_float,
        init: *const c_float,
        seed: u64,
        out: *mut c_float,
    ) -> c_int;

    /// Allocation-free `evoc_hip_embed` on device buffers (static per-GPU scratch:
    /// n <= 65536, nnz <= 2097152, dim <= 16).
    pub fn evoc_hip_embed_noalloc(
        d_indptr: *const i64,
        d_indices: *const i32,
        d_weights: *const c_float,
        n: i64,
        nnz: i64,
        dim: i32,
        n_epochs: i32,
        noise_level: c_float,
        negative_sample_rate: c_float,
        initial_alpha: c_float,
        beta1: c_float,
        beta2: c_float,
        eps: c_float,
        random_init: i32,
        seed: u64,
        device_id: i32,
        d_emb: *mut c_float,
    ) -> c_int;

    pub fn evoc_hip_label_prop_init(
        indptr: *const i64,
        indices: *const i32,
        weights: *const c_float,
        n: i64,
        dim: i32,
        data: *const c_float,
        d: i64,
        seed: u64,
        out: *mut c_float,
    ) -> c_int;

    pub fn evoc_hip_core_sq(
        emb: *const c_float,
        n: i64,
        dim: i32,
        min_samples: i32,
        core_sq: *mut c_float,
    ) -> c_int;

    pub fn evoc_hip_mst(
        emb: *const c_float,
        n: i64,
        dim: i32,
        min_samples: i32,
        u: *mut i32,
        v: *mut i32,
        w: *mut c_float,
    ) -> c_int;

    pub fn evoc_hip_condense_tree(
        left: *const i64,
        right: *const i64,
        distance: *const c_float,
        size: *const i64,
        n: i64,
        min_cluster_size: i64,
        parent: *mut i64,
        child: *mut i64,
        lambda: *mut c_float,
        child_size: *mut i64,
        n_rows: *mut i64,
    ) -> c_int;

    pub fn evoc_hip_cluster_layers(
        emb: *const c_float,
        n: i64,
        dim: i32,
        min_samples: i32,
        base_min_cluster_size: i64,
        min_similarity_threshold: c_double,
        max_layers: i64,
        labels: *mut i64,
        strengths: *mut c_float,
        persistence: *mut c_double,
        n_layers: *mut i64,
    ) -> c_int;

    /// One forward/backward pass of the warm-start MLP (gradient checks).
    pub fn evoc_hip_mlp_loss_grads(
        x: *const c_float,
        t: *const c_float,
        rows: i32,
        d: i32,
        out: i32,
        w0: *const c_float,
        b0: *const c_float,
        w1: *const c_float,
        b1: *const c_float,
        w2: *const c_float,
        b2: *const c_float,
        gw0: *mut c_float,
        gb0: *mut c_float,
        gw1: *mut c_float,
        gb1: *mut c_float,
        gw2: *mut c_float,
        gb2: *mut c_float,
        y: *mut c_float,
        loss: *mut c_double,
    ) -> c_int;

    pub fn evoc_hip_search_n_clusters(
        emb: *const c_float,
        n: i64,
        dim: i32,
        min_samples: i32,
        target_k: i64,
        labels: *mut i64,
        strengths: *mut c_float,
    ) -> c_int;
}
