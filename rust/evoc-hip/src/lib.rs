//! `evoc_hip()`: evoc-rs's clustering pipeline on one MI355X, with the same inputs and
//! result as `evoc_rs::evoc` (src/lib.rs:199-212 of evoc-rs 0.2.7):
//!
//! ```ignore
//! let res = evoc_hip(data.as_ref(), "exhaustive".into(), None, &EvocParams::default(),
//!                    &nn_params, 42, 1)?;
//! let labels = res.best_labels();
//! ```
//!
//! The kNN graph is always the exact exhaustive one (ties by index, self excluded): what
//! every `ann_type` approximates. `ann_type` is therefore accepted for every name the
//! reference knows; an unknown name is reported and searched exactly, where the reference
//! falls back to KmKnn (nearest_neighbour_cpu.rs:240-246).
use std::ffi::CStr;
use std::ptr;

use evoc_hip_sys as sys;
use evoc_rs::prelude::{EvocErrors, NearestNeighbourParamsEvoc, PreComputedKnn};
use evoc_rs::{EvocParams, EvocResult};
use faer::MatRef;

/// Errors of `evoc_hip`: the reference's own (`EvocErrors`, src/errors.rs:6-12) and the
/// device library's status codes.
#[derive(Debug, thiserror::Error)]
pub enum EvocHipError {
    #[error(transparent)]
    Evoc(#[from] EvocErrors),
    #[error("invalid argument: {0}")]
    InvalidArgument(String),
    #[error("HIP error: {0}")]
    Hip(String),
    #[error("no usable HIP device: {0}")]
    NoDevice(String),
    #[error("internal error: {0}")]
    Internal(String),
}

/// Device settings that have no counterpart in `NearestNeighbourParamsEvoc`.
#[derive(Clone, Debug)]
pub struct HipOptions {
    /// `EVOC_KNN_AUTO` (bf16 MFMA when the input is bf16-representable, else the hi/lo
    /// split), `EVOC_KNN_SPLIT` or `EVOC_KNN_BF16`; all return identical neighbours.
    pub knn_precision: i32,
    /// Learned MLP warm start instead of the label-propagation layout (accepted only if it
    /// passes its check on the training subsample).
    pub warm_start: bool,
    /// Embedding epochs of a warm-started run (0: 30% of `n_epochs`).
    pub warm_epochs: i32,
    /// MLP training subsample (0: 50,000 points).
    pub mlp_subsample: i32,
    /// HIP device ordinal.
    pub device_id: i32,
}

impl Default for HipOptions {
    fn default() -> Self {
        Self { knn_precision: sys::EVOC_KNN_AUTO, warm_start: false, warm_epochs: 0,
               mlp_subsample: 0, device_id: 0 }
    }
}

const ANN_TYPES: &[&str] = &["annoy", "balltree", "exhaustive", "hnsw", "ivf", "kmknn", "nndescent"];

/// `evoc_rs::evoc` with every stage on the GPU.
pub fn evoc_hip(
    data: MatRef<f32>,
    ann_type: String,
    precomputed_knn: PreComputedKnn<f32>,
    evoc_params: &EvocParams<f32>,
    nn_params: &NearestNeighbourParamsEvoc<f32>,
    seed: usize,
    verbose: usize,
) -> Result<EvocResult<f32>, EvocHipError> {
    evoc_hip_with(data, ann_type, precomputed_knn, evoc_params, nn_params, &HipOptions::default(),
                  seed, verbose)
}

/// `evoc_hip` with explicit device settings.
#[allow(clippy::too_many_arguments)]
pub fn evoc_hip_with(
    data: MatRef<f32>,
    ann_type: String,
    precomputed_knn: PreComputedKnn<f32>,
    evoc_params: &EvocParams<f32>,
    nn_params: &NearestNeighbourParamsEvoc<f32>,
    options: &HipOptions,
    seed: usize,
    verbose: usize,
) -> Result<EvocResult<f32>, EvocHipError> {
    if verbose > 0 && !ANN_TYPES.contains(&ann_type.to_lowercase().as_str()) {
        println!("Unrecognised approximate nearest neighbour method provided: {ann_type:?}. \
                  The exact exhaustive search runs.");
    }
    let metric = match nn_params.dist_metric.to_lowercase().as_str() {
        "euclidean" => sys::EVOC_METRIC_EUCLIDEAN,
        "cosine" => sys::EVOC_METRIC_COSINE,
        other => return Err(EvocHipError::InvalidArgument(format!("unknown dist_metric {other:?}"))),
    };
    let params = c_params(evoc_params)?;
    let nn = sys::evoc_hip_nn_params_t {
        metric,
        knn_precision: options.knn_precision,
        warm_start: options.warm_start as i32,
        warm_epochs: options.warm_epochs,
        mlp_subsample: options.mlp_subsample,
        device_id: options.device_id,
    };

    // faer allows negative strides; the library takes non-negative ones, so such a view
    // is packed row-major first.
    let (n, d) = (data.nrows(), data.ncols());
    let packed: Vec<f32>;
    let (ptr_x, rs, cs) = if data.row_stride() >= 0 && data.col_stride() >= 0 {
        (data.as_ptr(), data.row_stride() as i64, data.col_stride() as i64)
    } else {
        packed = (0..n).flat_map(|i| (0..d).map(move |j| data[(i, j)])).collect();
        (packed.as_ptr(), d as i64, 1)
    };

    // PreComputedKnn (src/prelude.rs:85): rows may be ragged; short rows are padded with
    // index -1, which the library skips.
    let (pre_idx, pre_dist, pre_k) = match &precomputed_knn {
        Some((idx, dist)) => {
            if idx.len() != n || dist.len() != n {
                return Err(EvocHipError::InvalidArgument(
                    "precomputed kNN must have one row per data point".into()));
            }
            let k = idx.iter().map(|r| r.len()).max().unwrap_or(0);
            let mut fi = vec![-1i32; n * k];
            let mut fd = vec![f32::INFINITY; n * k];
            for i in 0..n {
                if idx[i].len() != dist[i].len() {
                    return Err(EvocHipError::InvalidArgument(format!(
                        "precomputed kNN row {i}: {} indices, {} distances",
                        idx[i].len(), dist[i].len())));
                }
                for (t, (&j, &dj)) in idx[i].iter().zip(&dist[i]).enumerate() {
                    fi[i * k + t] = i32::try_from(j).map_err(|_| {
                        EvocHipError::InvalidArgument(format!("precomputed index {j} too large"))
                    })?;
                    fd[i * k + t] = dj;
                }
            }
            (fi, fd, k as i32)
        }
        None => (Vec::new(), Vec::new(), 0),
    };

    let mut out: *mut sys::evoc_result_t = ptr::null_mut();
    let status = unsafe {
        sys::evoc_hip_run(
            ptr_x, n as i64, d as i64, rs, cs,
            if pre_k > 0 { pre_idx.as_ptr() } else { ptr::null() },
            if pre_k > 0 { pre_dist.as_ptr() } else { ptr::null() },
            pre_k, &params, &nn, seed as u64, verbose as i32, &mut out,
        )
    };
    check(status)?;
    let guard = ResultGuard(out);
    let r = unsafe { &*guard.0 };
    let (n, k, nl) = (r.n as usize, r.k as usize, r.n_layers as usize);
    let slice = |p: *const f32, len: usize| unsafe { std::slice::from_raw_parts(p, len) };
    let labels = unsafe { std::slice::from_raw_parts(r.cluster_layers, nl * n) };
    let strengths = slice(r.membership_strengths, nl * n);
    let pers = unsafe { std::slice::from_raw_parts(r.persistence_scores, nl) };
    let nn_i = unsafe { std::slice::from_raw_parts(r.nn_indices, n * k) };
    let nn_d = slice(r.nn_distances, n * k);
    Ok(EvocResult {
        cluster_layers: labels.chunks(n.max(1)).take(nl).map(|c| c.to_vec()).collect(),
        membership_strengths: strengths.chunks(n.max(1)).take(nl).map(|c| c.to_vec()).collect(),
        persistence_scores: pers.to_vec(),
        // rows shorter than k (k > n - 1, or -1 padding of a precomputed graph) end early
        nn_indices: (0..n)
            .map(|i| nn_i[i * k..(i + 1) * k].iter().filter(|&&j| j >= 0).map(|&j| j as usize).collect())
            .collect(),
        nn_distances: (0..n)
            .map(|i| (0..k).filter(|&t| nn_i[i * k + t] >= 0).map(|t| nn_d[i * k + t]).collect())
            .collect(),
    })
}

fn c_params(p: &EvocParams<f32>) -> Result<sys::evoc_params_t, EvocHipError> {
    let to_i32 = |v: usize, what: &str| {
        i32::try_from(v).map_err(|_| EvocHipError::InvalidArgument(format!("{what} too large")))
    };
    Ok(sys::evoc_params_t {
        n_neighbours: to_i32(p.n_neighbours, "n_neighbours")?,
        noise_level: p.noise_level,
        n_epochs: to_i32(p.n_epochs, "n_epochs")?,
        embedding_dim: to_i32(p.embedding_dim.unwrap_or(0), "embedding_dim")?,
        neighbour_scale: p.neighbour_scale,
        symmetrise: p.symmetrise as i32,
        min_samples: to_i32(p.min_samples, "min_samples")?,
        base_min_cluster_size: to_i32(p.base_min_cluster_size, "base_min_cluster_size")?,
        approx_n_clusters: to_i32(p.approx_n_clusters.unwrap_or(0), "approx_n_clusters")?,
        min_similarity_threshold: p.min_similarity_threshold,
        max_layers: to_i32(p.max_layers, "max_layers")?,
    })
}

fn check(status: i32) -> Result<(), EvocHipError> {
    if status == sys::EVOC_OK {
        return Ok(());
    }
    let msg = unsafe {
        let p = sys::evoc_hip_last_error();
        if p.is_null() { String::new() } else { CStr::from_ptr(p).to_string_lossy().into_owned() }
    };
    Err(match status {
        sys::EVOC_ERR_INVALID_ARGUMENT => EvocHipError::InvalidArgument(msg),
        sys::EVOC_ERR_HIP => EvocHipError::Hip(msg),
        sys::EVOC_ERR_NO_DEVICE => EvocHipError::NoDevice(msg),
        _ => EvocHipError::Internal(msg),
    })
}

/// Frees the library's result when the Rust copy has been made (or on an early return).
struct ResultGuard(*mut sys::evoc_result_t);

impl Drop for ResultGuard {
    fn drop(&mut self) {
        if !self.0.is_null() {
            unsafe { sys::evoc_hip_result_free(self.0) };
        }
    }
}
