classdef GPEtaskHIP < GPEtask
    % GPEtaskHIP - GPEtask whose split-step solvers run on an MI355X.
    %
    % Same properties, methods and side effects as GPEtask (lib/@GPEtask/GPEtask.m);
    % solve_split and groundstate_itp call libgpehip.so through the MEX gateway
    % gpe_mex (gpe-matlab_amd/mex/gpe_mex.cpp).  The state stays on the GPU between
    % outer iterations; it is copied back for a user_callback or show_image, and once
    % at the end of a run.  The grid must come from grid2d(xmax,nx,ymax,ny) or
    % grid3d(xmax,nx,ymax,ny,zmax,nz) (symmetric linspace axes), with supported
    % lengths 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384 or 512.
    %
    %   grid = grid3d(12,256,12,256,12,256);
    %   task = GPEtaskHIP(grid, @(X,Y,Z) 0.5*(X.^2+Y.^2+16*Z.^2));
    %   task.Ntotal = 1e6; task.g = 0.02; task.gamma = 0.01;
    %   task.init_state = phi0;
    %   phi = task.solve_split(0.002, 50, 20);

    properties (SetAccess = protected)
        hip = uint64(0)   % gpe_mex task handle
    end

    methods
        function obj = GPEtaskHIP(grid, trappot)
            obj = obj@GPEtask(grid, trappot);
            obj.hip = gpe_mex('create', hip_grid_sizes(grid), hip_grid_extents(grid));
            if obj.fftw_set_wisdom
                % the counterpart of fftw('planner','measure') (GPEtask.m:55-58)
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
        [phi, varargout] = groundstate_itp(task, dt, eps, phi0)
    end
end

function n = hip_grid_sizes(grid)
    if grid.ndims == 3
        n = [grid.nx grid.ny grid.nz];
    else
        n = [grid.nx grid.ny];
    end
end

function xm = hip_grid_extents(grid)
    % the library builds every axis as linspace(-xmax, xmax, n) (grid2d/init.m:11-12)
    ax = {grid.x, grid.y};
    if grid.ndims == 3
        ax{3} = grid.z;
    end
    xm = zeros(1, numel(ax));
    for a = 1:numel(ax)
        v = ax{a};
        xm(a) = v(end);
        if max(abs(v - linspace(-v(end), v(end), numel(v)))) > 0
            error('GPEtaskHIP:grid', 'axis %d is not linspace(-xmax, xmax, n)', a);
        end
    end
end
