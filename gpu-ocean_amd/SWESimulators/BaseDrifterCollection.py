"""Host-side logic shared by drifter collections
(gpu_ocean/SWESimulators/BaseDrifterCollection.py:33-425): the observation
slot after the drifters, periodic-aware distances and innovations, Gaussian
and Cauchy weights, the collection mean, resampling and the initial
placements.  These operate on downloaded positions (a few values per
drifter); the device work is in GPUDrifterCollection."""

from __future__ import annotations

import abc

import numpy as np

from . import Common


class BaseDrifterCollection(object):
    """Abstract collection of drifters plus one observation position."""

    __metaclass__ = abc.ABCMeta

    def __init__(self, numDrifters, observation_variance=0.1,
                 boundaryConditions=Common.BoundaryConditions(),
                 domain_size_x=1.0, domain_size_y=1.0):
        self.numDrifters = numDrifters
        self.obs_index = self.numDrifters          # the observation is the last slot
        self.observation_variance = observation_variance
        self.positions = None
        self.domain_size_x = domain_size_x
        self.domain_size_y = domain_size_y
        self.boundaryConditions = boundaryConditions

    def __del__(self):
        self.cleanUp()

    def copy(self):
        pass

    @abc.abstractmethod
    def getDrifterPositions(self):
        pass

    @abc.abstractmethod
    def getObservationPosition(self):
        pass

    @abc.abstractmethod
    def setDrifterPositions(self, newDrifterPositions):
        pass

    @abc.abstractmethod
    def setObservationPosition(self, newObservationPosition):
        pass

    @abc.abstractmethod
    def cleanUp(self):
        pass

    @abc.abstractmethod
    def enforceBoundaryConditions(self):
        pass

    def getNumDrifters(self):
        return self.numDrifters

    def getObservationVariance(self):
        return self.observation_variance

    def getBoundaryConditions(self):
        return self.boundaryConditions

    def getDomainSizeX(self):
        return self.domain_size_x

    def getDomainSizeY(self):
        return self.domain_size_y

    def setBoundaryConditions(self, boundaryConditions):
        self.boundaryConditions = boundaryConditions

    def setDomainSize(self, size_x, size_y):
        self.domain_size_x = size_x
        self.domain_size_y = size_y

    def uniformly_distribute_drifters(self, initialization_cov_drifters=None):
        """Drifters at the centres of a near-square grid of sub-domains
        (round(sqrt(n)) rows), then perturbed with the given covariance
        (BaseDrifterCollection.py:141-172)."""
        n = self.numDrifters
        rows = int(np.round(np.sqrt(n)))
        cols = int(np.ceil(1.0 * n / rows)) if rows > 0 else 0
        k = np.arange(n)
        pos = np.empty((n, 2))
        pos[:, 0] = (k % cols + 0.5) * self.domain_size_x / cols if n else 0
        pos[:, 1] = (k // cols + 0.5) * self.domain_size_y / rows if n else 0
        if initialization_cov_drifters is None:
            cov = np.zeros((2, 2))
        elif np.isscalar(initialization_cov_drifters):
            cov = np.eye(2) * initialization_cov_drifters
        else:
            cov = np.asarray(initialization_cov_drifters)
        assert cov.shape == (2, 2), 'initialization_cov_drifters has the wrong shape: ' + str(cov)
        for d in range(n):
            pos[d, :] = np.random.multivariate_normal(pos[d, :], cov)
        self.setDrifterPositions(pos)

    def _getClosestPositions(self, obs=None):
        """Each drifter's periodic image closest to the observation, per
        periodic direction (BaseDrifterCollection.py:174-203)."""
        pos = np.array(self.getDrifterPositions(), dtype=np.float64)
        bc = self.boundaryConditions
        if not (bc.isPeriodicNorthSouth() or bc.isPeriodicEastWest()):
            return pos
        o = self.getObservationPosition() if obs is None else obs
        for axis, periodic, L in ((0, bc.isPeriodicEastWest(), self.getDomainSizeX()),
                                  (1, bc.isPeriodicNorthSouth(), self.getDomainSizeY())):
            if periodic:
                cand = pos[:, axis][:, None] + np.array([-L, 0.0, L])[None, :]
                best = np.argmin(np.abs(cand - o[axis]), axis=1)
                pos[:, axis] = cand[np.arange(len(pos)), best]
        return pos

    def getInnovations(self, obs=None):
        """d = y - H(x) per drifter, periodic images considered
        (BaseDrifterCollection.py:220-238)."""
        if obs is None:
            obs = self.getObservationPosition()
            closest = self._getClosestPositions()
        else:
            closest = self._getClosestPositions(obs)
        return np.asarray(obs, dtype=np.float64)[None, :] - closest

    def getDistances(self, obs=None):
        inn = self.getInnovations(obs)
        return np.sqrt(inn[:, 0] ** 2 + inn[:, 1] ** 2)

    def getGaussianWeight(self, distances=None, normalize=True):
        if distances is None:
            distances = self.getDistances()
        var = self.getObservationVariance()
        w = (1.0 / np.sqrt(2 * np.pi * var)) * np.exp(-(distances ** 2 / (2 * var)))
        return w / np.sum(w) if normalize else w

    def getCauchyWeight(self, distances=None, normalize=True):
        if distances is None:
            distances = self.getDistances()
        var = self.getObservationVariance()
        w = 1.0 / (np.pi * np.sqrt(var) * (1 + (distances ** 2 / var)))
        return w / np.sum(w) if normalize else w

    def _enforceBoundaryConditionsOnPosition(self, x, y):
        """One periodic wrap of (x, y); a negative y is wrapped by the x size,
        as in the reference (BaseDrifterCollection.py:282-296)."""
        if x < 0:
            x = self.domain_size_x + x
        if y < 0:
            y = self.domain_size_x + y
        if x > self.domain_size_x:
            x = x - self.domain_size_x
        if y > self.domain_size_y:
            y = y - self.domain_size_y
        return x, y

    def getCollectionMean(self, obs=None):
        mean_x, mean_y = np.mean(self._getClosestPositions(obs=obs), axis=0)
        mean_x, mean_y = self._enforceBoundaryConditionsOnPosition(mean_x, mean_y)
        return np.array([mean_x, mean_y])

    def resample(self, newSampleIndices, reinitialization_variance):
        """Positions of the selected drifters; duplicates after the first are
        drawn around it with the given variance (BaseDrifterCollection.py:311-348)."""
        old = np.array(self.getDrifterPositions(), dtype=np.float64)
        if self.getNumDrifters() != len(newSampleIndices):
            raise RuntimeError("ERROR: The size of the new ensemble differs from the old size!\n" +
                               "(old size, new size): " + str((self.getNumDrifters(), len(newSampleIndices))) +
                               "\nWe can fix this in the future by requiring a function resizeEnsemble")
        new = np.zeros((len(newSampleIndices), 2))
        if reinitialization_variance == 0:
            new[:, :] = old[newSampleIndices, :]
        else:
            seen = np.zeros(self.getNumDrifters(), dtype=bool)
            var = np.eye(2) * reinitialization_variance
            for i, index in enumerate(newSampleIndices):
                if seen[index]:
                    new[i, :] = np.random.multivariate_normal(old[index, :], var)
                else:
                    new[i, :] = old[index, :]
                    seen[index] = True
        self.setDrifterPositions(new)
        self.enforceBoundaryConditions()

    def initializeUniform(self):
        """Uniform positions in the domain; the observation in the central
        half (BaseDrifterCollection.py:352-368)."""
        p = np.random.rand(self.getNumDrifters() + 1, 2)
        p[-1, :] = p[-1, :] * 0.5 + 0.25
        p[:, 0] *= self.domain_size_x
        p[:, 1] *= self.domain_size_y
        self.setDrifterPositions(p[:-1, :])
        self.setObservationPosition(p[-1, :])
