;; The hip twin of the reference's uncomplicate.bayadera.opencl
;; (src/clojure/uncomplicate/bayadera/opencl.clj): binds the MI355X engine
;; factory and the model source library for a body of code, so that
;; uncomplicate.bayadera.core (distribution, likelihood, posterior, sample,
;; histogram, mean, sd) and uncomplicate.bayadera.mcmc (init!, burn-in!,
;; anneal!, run-sampler!, acc-rate!, mix!) run unchanged on it.  The model
;; sources are the reference's own C99 sources (its device library's .cl
;; files), which the engine compiles with hipRTC through a prelude supplying
;; REAL, native_exp, native_log, pown, lgamma and NAN.
(ns uncomplicate.bayadera.hip
  (:require [uncomplicate.bayadera
             [core :refer [with-bayadera *bayadera-factory*]]
             [library :refer [with-library]]]
            [uncomplicate.bayadera.internal.device
             [models :as models]
             [amd-cdna4 :refer [hip-bayadera-factory]]]))

(def source-library (models/source-library "uncomplicate/bayadera/internal/device/opencl/%s.cl"))

(def device-library (partial models/device-library source-library))

(defmacro with-default-library [factory & body]
  `(with-library (device-library ~factory)
     ~@body))

(defmacro with-default-bayadera
  "Device 0 with WGS 256, as the reference's tests run (amd_gcn_test.clj:33)."
  [& body]
  `(with-bayadera hip-bayadera-factory [0 256]
     (with-default-library *bayadera-factory*
       ~@body)))
