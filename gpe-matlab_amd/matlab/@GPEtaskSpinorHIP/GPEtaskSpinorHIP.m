classdef GPEtaskSpinorHIP < GPEtaskSpinor
    % GPEtaskSpinorHIP - GPEtaskSpinor whose solve_split runs on an MI355X.
    %
    % Same properties and side effects as GPEtaskSpinor
    % (lib/@GPEtaskSpinor/GPEtaskSpinor.m); solve_split calls libgpehip.so through the
    % MEX gateway gpe_mex with the fused two-component step (phase and the 2x2 Rabi
    % rotation in one x sweep).  Grid requirements as for GPEtaskHIP.

    properties (SetAccess = protected)
        hip = uint64(0)   % gpe_mex task handle
    end

    methods
        function obj = GPEtaskSpinorHIP(grid, trappot)
            obj = obj@GPEtaskSpinor(grid, trappot);
            if grid.ndims == 3
                n = [grid.nx grid.ny grid.nz];
                xm = [grid.x(end) grid.y(end) grid.z(end)];
            else
                n = [grid.nx grid.ny];
                xm = [grid.x(end) grid.y(end)];
            end
            obj.hip = gpe_mex('create', n, xm);
            if obj.fftw_set_wisdom
                gpe_mex('tune', obj.hip);
            end
        end

        function delete(obj)
            if obj.hip ~= 0
                gpe_mex('destroy', obj.hip);
                obj.hip = uint64(0);
            end
        end

        phi = solve_split(task, ddt, niter_inner, niter_outer)
    end
end
