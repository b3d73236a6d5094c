;; The MI355X (gfx950) engine behind Bayadera's protocols: every engine type
;; of the reference's OpenCL engine namespace
;; (src/clojure/uncomplicate/bayadera/internal/device/amd_gcn.clj) as a
;; deftype over the C-ABI of libbayadera_hip.so (include/bayadera_hip.h),
;; called through JNA.  Host values stay neanderthal native-float vectors and
;; column-major ge matrices; the library copies them to HBM at the call and
;; back, so no device buffer crosses this namespace.  Each engine holds an
;; opaque handle and releases it in Releaseable.release.  Handles are not
;; thread-safe, as the reference's engines are not (amd_gcn.clj:273-274).
;;
;; Not run in this repository (no JVM in its image): the same C-ABI is
;; exercised by bayadera_amd/engine.py (ctypes) in the test suite, which this
;; namespace mirrors call for call.
(ns uncomplicate.bayadera.internal.device.amd-cdna4
  (:require [clojure.string]
            [uncomplicate.commons.core :refer [Releaseable release let-release Info]]
            [uncomplicate.commons.utils :refer [dragan-says-ex]]
            [uncomplicate.neanderthal
             [core :refer [vctr ge ncols mrows dim]]
             [block :refer [buffer offset stride]]
             [native :refer [native-float]]
             [vect-math :refer [sqrt!]]]
            [uncomplicate.neanderthal.internal.api :as na]
            [uncomplicate.bayadera.internal.protocols :refer :all])
  (:import [com.sun.jna Function Native NativeLibrary Pointer Memory]
           [com.sun.jna.ptr PointerByReference DoubleByReference IntByReference
            LongByReference]
           [java.nio ByteBuffer]))

;; ============================ C-ABI plumbing =================================

(def ^:private ^NativeLibrary lib (NativeLibrary/getInstance "bayadera_hip"))

(def ^:const BH_ERR_ARG -1)
(def ^:const BH_ERR_RANGE -4)
(def ^:const BH_PARTNER_FLAT 0)
(def ^:const BH_PARTNER_WALKER 1)
(def ^:const BH_GLM_LOGISTIC 1)
(def ^:const BH_GLM_GAUSSIAN 2)
(def ^:const BH_GLM_FP64 0)
(def ^:const BH_GLM_FP32 1)
(def ^:const BH_GLM_BF16 2)
(def ^:const BH_GLM_BF16X3 3)
(def ^:const BH_GLM_FP16X2 4)
(def ^:const BH_DIRECT_KINDS {"uniform" 0 "gaussian" 1 "exponential" 2 "erlang" 3})

(defn- last-error ^String []
  (.invokeString (.getFunction lib "bh_last_error") (object-array 0) false))

(defn- call!
  "Calls a C-ABI function; a failure becomes the reference's exception kind:
  IllegalArgumentException for argument and size rules (BH_ERR_ARG,
  BH_ERR_RANGE: the walker-count rule, a short acor series), ex-info with the
  status code otherwise (a HIP error, a model that fails to compile)."
  [^String fname & args]
  (let [rc (.invokeInt (.getFunction lib fname) (object-array args))]
    (when-not (zero? rc)
      (let [msg (last-error)]
        (if (or (= rc BH_ERR_ARG) (= rc BH_ERR_RANGE))
          (throw (IllegalArgumentException. msg))
          (throw (ex-info msg {:code rc :fn fname})))))
    rc))

(defn- ptr
  "The native address of a neanderthal native-float vector or ge matrix
  (a direct ByteBuffer), at its offset."
  ^Pointer [x]
  (let [^ByteBuffer b (buffer x)]
    (.share (Native/getDirectBufferPointer b) (* Float/BYTES (long (offset x))))))

(defn- column-major! [x]
  (when-not (= (mrows x) (stride x))
    (dragan-says-ex "The engine takes contiguous column-major matrices."
                    {:mrows (mrows x) :stride (stride x)}))
  x)

(defn- handle ^Pointer [^PointerByReference out] (.getValue out))

(defn- build-log ^Memory [] (doto (Memory. 65536) (.clear)))

(defn- model-source ^String [model]
  (clojure.string/join "\n" (source model)))

;; ============================ Direct sampler =================================

(deftype HipDirectSamplerEngine [^Pointer ctx ^long kind]
  Releaseable
  (release [_] true)
  RandomSamplerEngine
  (sample [_ seed params res]
    (if (and (zero? (rem (ncols res) 4)) (zero? (rem (offset res) 4)))
      (do (call! "bh_direct_sample" ctx (int kind) (int seed) (ptr params) (int (dim params))
                 (ptr res) (long (ncols res)))
          res)
      (dragan-says-ex "GCN direct sampler supports only matrices with ncols and offset that are multiple of 4."
                      {:ncols (ncols res) :offset (offset res)}))))

;; ============================ Distribution engine ============================

(deftype HipDistributionEngine [^Pointer h dist-model]
  Releaseable
  (release [_] (call! "bh_dist_engine_destroy" h) true)
  ModelProvider
  (model [_] dist-model)
  DensityEngine
  (log-density [_ params x]
    (let-release [res (vctr (native-float) (ncols x))]
      (call! "bh_log_density" h (ptr params) (int (dim params)) (ptr (column-major! x))
             (int (mrows x)) (long (ncols x)) (ptr res))
      res))
  (density [_ params x]
    (let-release [res (vctr (native-float) (ncols x))]
      (call! "bh_density" h (ptr params) (int (dim params)) (ptr (column-major! x))
             (int (mrows x)) (long (ncols x)) (ptr res))
      res)))

;; ============================ Likelihood engine ==============================

(deftype HipLikelihoodEngine [^Pointer h lik-model]
  Releaseable
  (release [_] (call! "bh_lik_engine_destroy" h) true)
  ModelProvider
  (model [_] lik-model)
  DensityEngine
  (log-density [_ data x]
    (let-release [res (vctr (native-float) (ncols x))]
      (call! "bh_loglik" h (ptr data) (int (dim data)) (ptr (column-major! x)) (int (mrows x))
             (long (ncols x)) (ptr res))
      res))
  (density [_ data x]
    (let-release [res (vctr (native-float) (ncols x))]
      (call! "bh_lik" h (ptr data) (int (dim data)) (ptr (column-major! x)) (int (mrows x))
             (long (ncols x)) (ptr res))
      res))
  LikelihoodEngine
  (evidence [_ data x]
    (let [r (DoubleByReference.)]
      (call! "bh_evidence" h (ptr data) (int (dim data)) (ptr (column-major! x)) (int (mrows x))
             (long (ncols x)) r)
      (.getValue r))))

;; ============================ Dataset engine =================================

(deftype HipDatasetEngine [^Pointer ctx ^long WGS]
  Releaseable
  (release [_] true)
  DatasetEngine
  (data-mean [_ data-matrix]
    (let-release [res (vctr (native-float) (mrows data-matrix))]
      (call! "bh_dataset_mean" ctx (ptr data-matrix) (int (mrows data-matrix))
             (long (ncols data-matrix)) (int (stride data-matrix)) (ptr res))
      res))
  (data-variance [_ data-matrix]
    (let-release [res (vctr (native-float) (mrows data-matrix))]
      (call! "bh_dataset_variance" ctx (ptr data-matrix) (int (mrows data-matrix))
             (long (ncols data-matrix)) (int (stride data-matrix)) (ptr res))
      res))
  EstimateEngine
  (histogram [_ data-matrix]
    (let [m (mrows data-matrix)]
      (let-release [limits (ge (native-float) 2 m)
                    pdf (ge (native-float) WGS m)
                    bin-ranks (ge (native-float) WGS m)]
        (call! "bh_dataset_histogram" ctx (ptr data-matrix) (int m) (long (ncols data-matrix))
               (int (stride data-matrix)) (ptr limits) (ptr pdf) (ptr bin-ranks))
        (->Histogram limits pdf bin-ranks)))))

;; ============================ Autocorrelation ================================

(deftype HipAcorEngine [^Pointer ctx]
  Releaseable
  (release [_] true)
  AcorEngine
  (acor [_ data-matrix]
    (let [d (mrows data-matrix) n (ncols data-matrix) lag (IntByReference.)]
      (let-release [tau (vctr (native-float) d)
                    mean (vctr (native-float) d)
                    sigma (vctr (native-float) d)]
        (call! "bh_acor" ctx (ptr (column-major! data-matrix)) (int d) (int n) (ptr tau)
               (ptr mean) (ptr sigma) lag)
        (->Autocorrelation tau mean sigma n (.getValue lag))))))

;; ======================== MCMC engine ========================================

(declare ->HipStretch)

(deftype HipStretch [^Pointer h ^long DIM ^long WGS ^long walker-count model]
  Releaseable
  (release [_] (call! "bh_stretch_destroy" h) true)
  Info
  (info [_]
    (let [w (IntByReference.) it (LongByReference.)]
      (call! "bh_stretch_info" h w it)
      {:walker-count (.getValue w) :iteration-counter (.getValue it)}))
  ModelProvider
  (model [_] model)
  MCMCStretch
  (init-move! [this _ a]
    ;; the engine keeps the means accumulator itself (one step reserved,
    ;; as acc-rate! uses it); run-sampler! reserves its n steps
    (call! "bh_stretch_init_move" h (float a) 1)
    this)
  (move! [this] (call! "bh_stretch_move" h) this)
  (move-bare! [this] (call! "bh_stretch_move_bare" h) this)
  (set-temperature! [this t] (call! "bh_stretch_set_temperature" h (double t)) this)
  RandomSampler
  (sample! [this] (sample! this walker-count))
  (sample! [_ n-or-res]
    (let-release [res (if (integer? n-or-res)
                        (ge (native-float) DIM n-or-res {:raw true})
                        (column-major! n-or-res))]
      (call! "bh_stretch_sample" h (ptr res) (long (ncols res)))
      res))
  MCMC
  (init! [this seed] (call! "bh_stretch_init" h (int seed)) this)
  (init-position! [this position]
    (call! "bh_stretch_init_position_from" h (.h ^HipStretch position))
    this)
  (init-position! [this seed limits]
    (call! "bh_stretch_init_position" h (int seed) (ptr (column-major! limits)))
    this)
  (burn-in! [this n a] (call! "bh_stretch_burn_in" h (long n) (float a)) this)
  (anneal! [this schedule n a]
    ;; the schedule is a Clojure function: its temperatures are computed here
    ;; and passed as an array (the library has no callbacks)
    (let [temps (double-array (map schedule (range n)))]
      (call! "bh_stretch_anneal" h temps (long n) (float a))
      this))
  (run-sampler! [_ n a]
    (let [acc (DoubleByReference.) lag (IntByReference.)]
      (let-release [tau (vctr (native-float) DIM)
                    mean (vctr (native-float) DIM)
                    sigma (vctr (native-float) DIM)]
        (call! "bh_stretch_run" h (long n) (float a) acc nil (ptr tau) (ptr mean) (ptr sigma) lag)
        {:acceptance-rate (.getValue acc)
         :a (float a)
         :autocorrelation (->Autocorrelation tau mean sigma n (.getValue lag))})))
  (acc-rate! [_ a]
    (let [r (DoubleByReference.)]
      (call! "bh_stretch_acc_rate" h (float a) r)
      (.getValue r)))
  EstimateEngine
  (histogram [this] (histogram! this 1))
  (histogram! [_ cycles]
    (let-release [limits (ge (native-float) 2 DIM)
                  pdf (ge (native-float) WGS DIM)
                  bin-ranks (ge (native-float) WGS DIM)]
      (call! "bh_stretch_histogram" h (int cycles) (ptr limits) (ptr pdf) (ptr bin-ranks))
      (->Histogram limits pdf bin-ranks)))
  Location
  (mean [_]
    (let-release [res (vctr (native-float) DIM)]
      (call! "bh_stretch_mean" h (ptr res))
      res))
  Spread
  (variance [_]
    (let-release [res (vctr (native-float) DIM)]
      (call! "bh_stretch_variance" h (ptr res))
      res))
  (sd [this] (sqrt! (variance this))))

(defn- sampler-options!
  "Engine options without a reference counterpart, from the sampler's option
  map: :partner (:flat, the reference's, or :walker), and for GLM samplers
  :precision (:fp64 :fp32 :bf16 :bf16x3 :fp16x2) and :guard (log units)."
  [^Pointer h {:keys [partner precision guard]}]
  (when partner
    (call! "bh_stretch_set_partner" h (int (if (= :walker partner) BH_PARTNER_WALKER
                                               BH_PARTNER_FLAT))))
  (when precision
    (call! "bh_stretch_set_glm_precision" h
           (int ({:fp64 BH_GLM_FP64 :fp32 BH_GLM_FP32 :bf16 BH_GLM_BF16
                  :bf16x3 BH_GLM_BF16X3 :fp16x2 BH_GLM_FP16X2} precision))))
  (when guard
    (call! "bh_stretch_set_glm_guard" h (double guard)))
  h)

(deftype HipStretchFactory [^Pointer f model ^long DIM ^long WGS options]
  Releaseable
  (release [_] (call! "bh_stretch_factory_destroy" f) true)
  ModelProvider
  (model [_] model)
  SamplerFactory
  (create-sampler [_ seed walker-count params]
    ;; the walker-count rule (a multiple of 2 WGS, amd_gcn.clj:516, 566-567)
    ;; is enforced by bh_stretch_create: BH_ERR_RANGE -> IllegalArgumentException
    (let [out (PointerByReference.)
          p (when params (ptr params))]
      (call! "bh_stretch_create" f (int seed) (int walker-count) p
             (int (if params (dim params) 0)) out)
      (let-release [s (->HipStretch (sampler-options! (handle out) options) DIM WGS
                                    (long walker-count) model)]
        s))))

;; ======================== GLM models declared by form ========================

(defn glm-stretch-factory
  "A sampler factory for a GLM posterior declared by form (no reference
  counterpart: the reference runs it as a user model whose LOGFN loops over
  the data rows).  model-map: {:glm :logistic :X X-bf16-bytes :N N :d d :y y
  :prior-sd s} or {:glm :gaussian :X X (N x d ge) :y y :beta-sd sb
  :log-sigma-sd ss}; the likelihood of a half-ensemble runs as one
  data-parallel pass (matrix cores for the logistic reduced precisions)."
  [^Pointer ctx {:keys [glm X N d y prior-sd beta-sd log-sigma-sd] :as model-map} WGS options]
  (let [out (PointerByReference.) log (build-log)]
    (case glm
      :logistic (call! "bh_glm_stretch_factory_create" ctx (int BH_GLM_LOGISTIC) X (ptr y)
                       (long N) (int d) (float prior-sd) out log (long (.size log)))
      :gaussian (call! "bh_glm_gaussian_factory_create" ctx (ptr X) (ptr y) (long (mrows X))
                       (int (ncols X)) (float beta-sd) (float log-sigma-sd) out log
                       (long (.size log))))
    (->HipStretchFactory (handle out) model-map (if (= :gaussian glm) (inc (long (ncols X))) d)
                         WGS options)))

;; ======================== Batched posteriors (new) ===========================

(defn run-batch!
  "Many independent posteriors of one user model in one launch per call (no
  reference counterpart): problems' params as one vector with offsets,
  their dimensions and hyperparameter counts; burn-in then run-sampler! with
  seeds seed0 + 2p, + 1 for problem p.  Returns one result map per problem."
  [^Pointer ctx model {:keys [params offsets dims hyper-len walkers seed0 burn-in steps a
                              limits]}]
  (let [out (PointerByReference.) log (build-log)
        P (count dims) max-dim (long (dimension model))]
    (call! "bh_batch_create" ctx (model-source model) (mcmc-logpdf model) (int max-dim) (int P)
           (int-array dims) (long-array offsets) (int-array hyper-len) (ptr params)
           (int walkers) out log (long (.size log)))
    (let [b (handle out)
          acc (double-array P) mean (double-array (* P max-dim)) sd (double-array (* P max-dim))]
      (try
        (call! "bh_batch_init_position" b (int seed0) (ptr limits))
        (call! "bh_batch_burn_in" b (int seed0) (long burn-in) (float a))
        (call! "bh_batch_run" b (int seed0) (long steps) (float a) acc mean sd)
        (mapv (fn [p] {:acceptance-rate (aget acc p)
                       :mean (java.util.Arrays/copyOfRange mean (* p max-dim)
                                                           (+ (* p max-dim) (long (nth dims p))))
                       :sd (java.util.Arrays/copyOfRange sd (* p max-dim)
                                                         (+ (* p max-dim) (long (nth dims p))))})
              (range P))
        (finally (call! "bh_batch_destroy" b))))))

;; =========================== Bayadera factory  ===============================

(defn- compile-engine [fname ^Pointer ctx & args]
  (let [out (PointerByReference.) log (build-log)]
    (apply call! fname ctx (concat args [out log (long (.size log))]))
    (handle out)))

(defrecord HipBayaderaFactory [^Pointer ctx ^long compute-units ^long WGS neanderthal-factory
                               dataset-eng acor-eng]
  Releaseable
  (release [_]
    (release neanderthal-factory)
    (call! "bh_ctx_destroy" ctx)
    true)
  na/MemoryContext
  (compatible? [_ o]
    (or (satisfies? DeviceModel o) (map? o) (na/compatible? neanderthal-factory o)))
  na/FactoryProvider
  (factory [_] neanderthal-factory)
  (native-factory [_] neanderthal-factory)
  EngineFactory
  (likelihood-engine [_ model]
    (->HipLikelihoodEngine (compile-engine "bh_lik_engine_create" ctx (model-source model)
                                           (loglik model))
                           model))
  (distribution-engine [_ model]
    (->HipDistributionEngine (compile-engine "bh_dist_engine_create" ctx (model-source model)
                                             (logpdf model) (int (params-size model)))
                             model))
  (dataset-engine [_] dataset-eng)
  (direct-sampler-engine [_ model]
    ;; the library's direct samplers (rng/NAME-sampler.cl) are built into the
    ;; engine; the model's logpdf name (NAME_logpdf) selects one
    (if-let [kind (and (seq (remove nil? (sampler-source model)))
                       (BH_DIRECT_KINDS (first (clojure.string/split (logpdf model) #"_"))))]
      (->HipDirectSamplerEngine ctx kind)
      (dragan-says-ex "No direct sampler for this model." {:logpdf (logpdf model)})))
  (mcmc-factory [this model]
    (if (and (map? model) (:glm model))
      (glm-stretch-factory ctx model WGS (:options model))
      (let [dim (long (dimension model))]
        (->HipStretchFactory (compile-engine "bh_stretch_factory_create" ctx (model-source model)
                                             (mcmc-logpdf model) (int dim)
                                             (int (params-size model)))
                             model dim WGS (when (map? model) (:options model))))))
  (processing-elements [_] (* compute-units WGS)))

(defn hip-bayadera-factory
  "The EngineFactory of one MI355X (gcn-bayadera-factory, amd_gcn.clj:666-678):
  device index and WGS (the reference tests use 256; the bin count of
  histograms and the walker-count quantum follow it)."
  ([^long device ^long WGS]
   (let [out (PointerByReference.)]
     (call! "bh_ctx_create" (int device) (int WGS) out)
     (let [ctx (handle out) cu (IntByReference.) wgs (IntByReference.)]
       (call! "bh_ctx_info" ctx cu wgs)
       (->HipBayaderaFactory ctx (.getValue cu) (.getValue wgs) (native-float)
                             (->HipDatasetEngine ctx (.getValue wgs)) (->HipAcorEngine ctx)))))
  ([] (hip-bayadera-factory 0 256)))
